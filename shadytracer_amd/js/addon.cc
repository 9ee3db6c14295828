// The Node N-API addon: libpt's C ABI (include/pt.h) as JavaScript functions, one to one. Structures cross as
// ArrayBuffers that the JS shim packs (raytrace_main.js); trace, setScene and foldParams return Promises settled on a
// worker thread once the GPU (or the compiler) has finished, so neither a trace nor a compile blocks the event loop
// (the reference's step() resolves on queue.onSubmittedWorkDone, lib/pathtracer/index.js:1005-1007; its setScene is
// async, raytrace_main.js:1145).
#include <node_api.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "pt.h"

namespace {

#define NAPI_OK(env, call)                                                              \
    do {                                                                                \
        if ((call) != napi_ok) {                                                        \
            napi_throw_error((env), nullptr, "N-API call failed: " #call);              \
            return nullptr;                                                             \
        }                                                                               \
    } while (0)

napi_value throwPt(napi_env env, pt_ctx* ctx, int code) {
    napi_value msg, err, codeV, restored;
    std::string text = std::string("libpt: ") + pt_last_error(ctx);
    napi_create_string_utf8(env, text.c_str(), text.size(), &msg);
    napi_create_error(env, nullptr, msg, &err);
    napi_create_int32(env, code, &codeV);
    napi_set_named_property(env, err, "code", codeV);
    napi_get_boolean(env, code == PT_ERR_COMPILE, &restored);
    napi_set_named_property(env, err, "restored", restored);
    napi_throw(env, err);
    return nullptr;
}

bool args(napi_env env, napi_callback_info info, size_t want, napi_value* out) {
    size_t n = want;
    if (napi_get_cb_info(env, info, &n, out, nullptr, nullptr) != napi_ok || n < want) {
        napi_throw_type_error(env, nullptr, "too few arguments");
        return false;
    }
    return true;
}

pt_ctx* ctxOf(napi_env env, napi_value v) {
    void* p = nullptr;
    if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
        napi_throw_type_error(env, nullptr, "not an engine handle (or destroyed)");
        return nullptr;
    }
    return *static_cast<pt_ctx**>(p);
}

bool bytesOf(napi_env env, napi_value v, void** data, size_t* len) {
    bool isAB = false, isTA = false;
    napi_is_arraybuffer(env, v, &isAB);
    if (isAB) return napi_get_arraybuffer_info(env, v, data, len) == napi_ok;
    napi_is_typedarray(env, v, &isTA);
    if (isTA) {
        napi_typedarray_type t;
        size_t n, off;
        napi_value ab;
        if (napi_get_typedarray_info(env, v, &t, &n, data, &ab, &off) != napi_ok) return false;
        size_t el = (t == napi_float64_array || t == napi_bigint64_array || t == napi_biguint64_array) ? 8
                  : (t == napi_float32_array || t == napi_int32_array || t == napi_uint32_array) ? 4
                  : (t == napi_int16_array || t == napi_uint16_array) ? 2 : 1;
        *len = n * el;
        return true;
    }
    napi_throw_type_error(env, nullptr, "expected an ArrayBuffer or a typed array");
    return false;
}

napi_value num(napi_env env, double v) { napi_value r; napi_create_double(env, v, &r); return r; }
napi_value undef(napi_env env) { napi_value r; napi_get_undefined(env, &r); return r; }

std::string str(napi_env env, napi_value v) {
    size_t n = 0;
    napi_valuetype t;
    napi_typeof(env, v, &t);
    if (t != napi_string) return std::string();
    napi_get_value_string_utf8(env, v, nullptr, 0, &n);
    std::string s(n, '\0');
    napi_get_value_string_utf8(env, v, &s[0], n + 1, &n);
    return s;
}

void finalizeHandle(napi_env, void* data, void*) {
    pt_ctx** box = static_cast<pt_ctx**>(data);
    if (*box) pt_destroy(*box);
    delete box;
}

// create(device, blocksPerCu, cacheDir) -> handle
napi_value Create(napi_env env, napi_callback_info info) {
    napi_value a[3];
    if (!args(env, info, 3, a)) return nullptr;
    int32_t dev = 0; uint32_t bpc = 0;
    napi_get_value_int32(env, a[0], &dev);
    napi_get_value_uint32(env, a[1], &bpc);
    std::string cache = str(env, a[2]);
    pt_opts o;
    memset(&o, 0, sizeof o);
    o.device = dev; o.blocks_per_cu = bpc; o.cache_dir = cache.empty() ? nullptr : cache.c_str();
    pt_ctx* ctx = nullptr;
    int rc = pt_create(&ctx, &o);
    if (rc) return throwPt(env, nullptr, rc);
    pt_ctx** box = new pt_ctx*(ctx);
    napi_value h;
    NAPI_OK(env, napi_create_external(env, box, finalizeHandle, nullptr, &h));
    return h;
}

napi_value Destroy(napi_env env, napi_callback_info info) {
    napi_value a[1];
    if (!args(env, info, 1, a)) return nullptr;
    void* p = nullptr;
    napi_get_value_external(env, a[0], &p);
    pt_ctx** box = static_cast<pt_ctx**>(p);
    if (box && *box) { pt_destroy(*box); *box = nullptr; }
    return undef(env);
}

napi_value AbiVersion(napi_env env, napi_callback_info) { return num(env, pt_abi_version()); }

napi_value StructSize(napi_env env, napi_callback_info info) {
    napi_value a[1];
    if (!args(env, info, 1, a)) return nullptr;
    int32_t w = 0;
    napi_get_value_int32(env, a[0], &w);
    return num(env, (double)pt_struct_size(w));
}

// compileScene(src, cacheDir) -> path (no GPU)
napi_value CompileScene(napi_env env, napi_callback_info info) {
    napi_value a[2];
    if (!args(env, info, 2, a)) return nullptr;
    std::string src = str(env, a[0]), cache = str(env, a[1]);
    char path[4096];
    int rc = pt_compile_scene(src.c_str(), cache.empty() ? nullptr : cache.c_str(), path, sizeof path);
    if (rc) return throwPt(env, nullptr, rc);
    napi_value r;
    napi_create_string_utf8(env, path, strlen(path), &r);
    return r;
}

// A scene build on a worker thread (the reference's setScene and its pipelines are asynchronous: hipRTC's compile,
// seconds for a large scene, never holds the event loop). Everything it reads is copied in first: the JS side may
// collect its buffers before the worker runs. The shim makes no other call on the context until the Promise settles.
struct BuildWork {
    napi_async_work work;
    napi_deferred deferred;
    pt_ctx* ctx;
    bool fold;                          // pt_fold_params instead of pt_set_scene
    std::string src;
    std::vector<float> params;
    std::vector<PtMaterial> mats;
    std::vector<PtLightFace> faces;
    uint32_t flags;
    int rc;
    std::string err;
    double ms;
};

void buildExecute(napi_env, void* data) {
    BuildWork* w = static_cast<BuildWork*>(data);
    const auto t0 = std::chrono::steady_clock::now();
    if (w->fold) {
        w->rc = pt_fold_params(w->ctx);
    } else {
        pt_scene_desc d;
        d.hip_src = w->src.c_str();
        d.params = w->params.data(); d.nparams = (uint32_t)w->params.size();
        d.materials = w->mats.data(); d.nmaterials = (uint32_t)w->mats.size();
        d.faces = w->faces.empty() ? nullptr : w->faces.data(); d.nfaces = (uint32_t)w->faces.size();
        d.flags = w->flags;
        w->rc = pt_set_scene(w->ctx, &d);
    }
    if (!w->rc) w->rc = pt_sync(w->ctx);
    if (w->rc) w->err = pt_last_error(w->ctx);
    w->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void buildComplete(napi_env env, napi_status, void* data) {
    BuildWork* w = static_cast<BuildWork*>(data);
    if (w->rc) {
        // the Error a failed build rejects with: code, and restored when the scene before is still the one loaded
        napi_value msg, err, codeV, restored;
        std::string text = "libpt: " + w->err;
        napi_create_string_utf8(env, text.c_str(), text.size(), &msg);
        napi_create_error(env, nullptr, msg, &err);
        napi_create_int32(env, w->rc, &codeV);
        napi_set_named_property(env, err, "code", codeV);
        napi_get_boolean(env, w->rc == PT_ERR_COMPILE, &restored);
        napi_set_named_property(env, err, "restored", restored);
        napi_reject_deferred(env, w->deferred, err);
    } else {
        napi_resolve_deferred(env, w->deferred, num(env, w->ms));
    }
    napi_delete_async_work(env, w->work);
    delete w;
}

napi_value queueBuild(napi_env env, BuildWork* w) {
    napi_value promise, name;
    if (napi_create_promise(env, &w->deferred, &promise) != napi_ok) { delete w; napi_throw_error(env, nullptr, "no promise"); return nullptr; }
    napi_create_string_utf8(env, w->fold ? "pt_fold_params" : "pt_set_scene", NAPI_AUTO_LENGTH, &name);
    NAPI_OK(env, napi_create_async_work(env, nullptr, name, buildExecute, buildComplete, w, &w->work));
    NAPI_OK(env, napi_queue_async_work(env, w->work));
    return promise;
}

// setScene(h, src, paramsF32, materialsBuf, nmat, facesBuf, nfaces, keepPicture) -> Promise of the build's ms
napi_value SetScene(napi_env env, napi_callback_info info) {
    napi_value a[8];
    if (!args(env, info, 7, a)) return nullptr;
    pt_ctx* ctx = ctxOf(env, a[0]);
    if (!ctx) return nullptr;
    void *pp, *mp, *fp; size_t pl, ml, fl;
    if (!bytesOf(env, a[2], &pp, &pl) || !bytesOf(env, a[3], &mp, &ml) || !bytesOf(env, a[5], &fp, &fl)) return nullptr;
    uint32_t nm = 0, nf = 0;
    napi_get_value_uint32(env, a[4], &nm);
    napi_get_value_uint32(env, a[6], &nf);
    if (ml < nm * sizeof(PtMaterial) || fl < nf * sizeof(PtLightFace)) {
        napi_throw_range_error(env, nullptr, "material or face buffer shorter than its count");
        return nullptr;
    }
    bool keep = false;
    size_t argc = 8;
    napi_value all[8];
    napi_get_cb_info(env, info, &argc, all, nullptr, nullptr);
    if (argc >= 8) napi_get_value_bool(env, all[7], &keep);
    BuildWork* w = new BuildWork();
    w->ctx = ctx;
    w->fold = false;
    w->src = str(env, a[1]);
    w->params.assign(static_cast<const float*>(pp), static_cast<const float*>(pp) + pl / 4);
    w->mats.assign(static_cast<const PtMaterial*>(mp), static_cast<const PtMaterial*>(mp) + nm);
    w->faces.assign(static_cast<const PtLightFace*>(fp), static_cast<const PtLightFace*>(fp) + nf);
    w->flags = keep ? PT_SCENE_KEEP_PICTURE : 0u;
    return queueBuild(env, w);
}

napi_value SetParams(napi_env env, napi_callback_info info) {
    napi_value a[2];
    if (!args(env, info, 2, a)) return nullptr;
    pt_ctx* ctx = ctxOf(env, a[0]);
    if (!ctx) return nullptr;
    void* p; size_t l;
    if (!bytesOf(env, a[1], &p, &l)) return nullptr;
    int rc = pt_set_params(ctx, static_cast<const float*>(p), (uint32_t)(l / 4));
    if (rc) return throwPt(env, ctx, rc);
    return undef(env);
}

// foldParams(h) -> Promise of the build's ms (a compile too)
napi_value FoldParams(napi_env env, napi_callback_info info) {
    napi_value a[1];
    if (!args(env, info, 1, a)) return nullptr;
    pt_ctx* ctx = ctxOf(env, a[0]);
    if (!ctx) return nullptr;
    BuildWork* w = new BuildWork();
    w->ctx = ctx;
    w->fold = true;
    w->flags = 0u;
    return queueBuild(env, w);
}

napi_value SetUniforms(napi_env env, napi_callback_info info) {
    napi_value a[2];
    if (!args(env, info, 2, a)) return nullptr;
    pt_ctx* ctx = ctxOf(env, a[0]);
    if (!ctx) return nullptr;
    void* p; size_t l;
    if (!bytesOf(env, a[1], &p, &l)) return nullptr;
    int rc = pt_set_uniforms(ctx, static_cast<const PtUniforms*>(p), l);
    if (rc) return throwPt(env, ctx, rc);
    return undef(env);
}

napi_value ClearAccum(napi_env env, napi_callback_info info) {
    napi_value a[1];
    if (!args(env, info, 1, a)) return nullptr;
    pt_ctx* ctx = ctxOf(env, a[0]);
    if (!ctx) return nullptr;
    int rc = pt_clear_accum(ctx);
    if (rc) return throwPt(env, ctx, rc);
    return undef(env);
}

struct TraceWork {
    napi_async_work work;
    napi_deferred deferred;
    pt_ctx* ctx;
    uint32_t samples;
    int rc;
    std::string err;
};

void traceExecute(napi_env, void* data) {
    TraceWork* w = static_cast<TraceWork*>(data);
    w->rc = pt_trace(w->ctx, w->samples);
    if (!w->rc) w->rc = pt_sync(w->ctx);
    if (w->rc) w->err = pt_last_error(w->ctx);
}

void traceComplete(napi_env env, napi_status, void* data) {
    TraceWork* w = static_cast<TraceWork*>(data);
    if (w->rc) {
        napi_value msg, err;
        std::string text = "libpt: " + w->err;
        napi_create_string_utf8(env, text.c_str(), text.size(), &msg);
        napi_create_error(env, nullptr, msg, &err);
        napi_reject_deferred(env, w->deferred, err);
    } else {
        napi_resolve_deferred(env, w->deferred, undef(env));
    }
    napi_delete_async_work(env, w->work);
    delete w;
}

// trace(h, samples) -> Promise, resolved once the GPU has run them
napi_value Trace(napi_env env, napi_callback_info info) {
    napi_value a[2];
    if (!args(env, info, 2, a)) return nullptr;
    pt_ctx* ctx = ctxOf(env, a[0]);
    if (!ctx) return nullptr;
    TraceWork* w = new TraceWork();
    w->ctx = ctx;
    napi_get_value_uint32(env, a[1], &w->samples);
    napi_value promise, name;
    NAPI_OK(env, napi_create_promise(env, &w->deferred, &promise));
    napi_create_string_utf8(env, "pt_trace", NAPI_AUTO_LENGTH, &name);
    NAPI_OK(env, napi_create_async_work(env, nullptr, name, traceExecute, traceComplete, w, &w->work));
    NAPI_OK(env, napi_queue_async_work(env, w->work));
    return promise;
}

// readAccum(h, f64accum | null, f64moments | null)
napi_value ReadAccum(napi_env env, napi_callback_info info) {
    napi_value a[3];
    if (!args(env, info, 3, a)) return nullptr;
    pt_ctx* ctx = ctxOf(env, a[0]);
    if (!ctx) return nullptr;
    void *ap = nullptr, *mp = nullptr; size_t al = 0, ml = 0;
    napi_valuetype t;
    napi_typeof(env, a[1], &t);
    if (t != napi_null && !bytesOf(env, a[1], &ap, &al)) return nullptr;
    napi_typeof(env, a[2], &t);
    if (t != napi_null && !bytesOf(env, a[2], &mp, &ml)) return nullptr;
    size_t npx = ap ? al / 32 : ml / 16;
    if (ap && mp && al / 32 != ml / 16) { napi_throw_range_error(env, nullptr, "accum and moments sizes disagree"); return nullptr; }
    int rc = pt_read_accum(ctx, static_cast<double*>(ap), static_cast<double*>(mp), npx);
    if (rc) return throwPt(env, ctx, rc);
    return undef(env);
}

// readStats(h) -> Float64Array of PT_S_WORDS counts
napi_value ReadStats(napi_env env, napi_callback_info info) {
    napi_value a[1];
    if (!args(env, info, 1, a)) return nullptr;
    pt_ctx* ctx = ctxOf(env, a[0]);
    if (!ctx) return nullptr;
    uint64_t s[PT_S_WORDS];
    int rc = pt_read_stats(ctx, s, PT_S_WORDS);
    if (rc) return throwPt(env, ctx, rc);
    napi_value ab, arr;
    void* data;
    NAPI_OK(env, napi_create_arraybuffer(env, PT_S_WORDS * 8, &data, &ab));
    for (int i = 0; i < PT_S_WORDS; ++i) static_cast<double*>(data)[i] = (double)s[i];
    NAPI_OK(env, napi_create_typedarray(env, napi_float64_array, PT_S_WORDS, ab, 0, &arr));
    return arr;
}

// autotune(h, samples) -> { blocksPerCu, refillMarch, refillOcclude, candidates, msPerSample, msBefore }
napi_value Autotune(napi_env env, napi_callback_info info) {
    napi_value a[2];
    if (!args(env, info, 2, a)) return nullptr;
    pt_ctx* ctx = ctxOf(env, a[0]);
    if (!ctx) return nullptr;
    uint32_t samples = 4;
    napi_get_value_uint32(env, a[1], &samples);
    PtTuning t;
    int rc = pt_autotune(ctx, samples, &t, sizeof t);
    if (rc) return throwPt(env, ctx, rc);
    napi_value o;
    NAPI_OK(env, napi_create_object(env, &o));
    napi_set_named_property(env, o, "blocksPerCu", num(env, t.blocksPerCu));
    napi_set_named_property(env, o, "refillMarch", num(env, t.refillMarch));
    napi_set_named_property(env, o, "refillOcclude", num(env, t.refillOcclude));
    napi_set_named_property(env, o, "candidates", num(env, t.candidates));
    napi_set_named_property(env, o, "msPerSample", num(env, t.msPerSample));
    napi_set_named_property(env, o, "msBefore", num(env, t.msBefore));
    return o;
}

// setTuning(h, blocksPerCu, refillMarch, refillOcclude)
napi_value SetTuning(napi_env env, napi_callback_info info) {
    napi_value a[4];
    if (!args(env, info, 4, a)) return nullptr;
    pt_ctx* ctx = ctxOf(env, a[0]);
    if (!ctx) return nullptr;
    PtTuning t;
    memset(&t, 0, sizeof t);
    napi_get_value_uint32(env, a[1], &t.blocksPerCu);
    napi_get_value_uint32(env, a[2], &t.refillMarch);
    napi_get_value_uint32(env, a[3], &t.refillOcclude);
    int rc = pt_set_tuning(ctx, &t, sizeof t);
    if (rc) return throwPt(env, ctx, rc);
    return undef(env);
}

// opticsTrace(device, scenes, n, results, far | null, lux | null) -> [walk ms, numbers ms]. scenes: n PtOpticsScene
// packed; results: room for n PtOpticsResult; far / lux: Float64Arrays of n x 181 x 360 / n x 256 x 256. One launch of
// ~30 ms for the bench's 64 scenes: run synchronously, as the optics page's run() awaits its whole batch.
napi_value OpticsTrace(napi_env env, napi_callback_info info) {
    napi_value a[6];
    if (!args(env, info, 6, a)) return nullptr;
    int32_t dev = 0; uint32_t n = 0;
    napi_get_value_int32(env, a[0], &dev);
    napi_get_value_uint32(env, a[2], &n);
    void *sp = nullptr, *rp = nullptr, *fp = nullptr, *lp = nullptr;
    size_t sl = 0, rl = 0, fl = 0, ll = 0;
    if (!bytesOf(env, a[1], &sp, &sl) || !bytesOf(env, a[3], &rp, &rl)) return nullptr;
    if (sl < (size_t)n * sizeof(PtOpticsScene) || rl < (size_t)n * sizeof(PtOpticsResult)) {
        napi_throw_range_error(env, nullptr, "scenes or results smaller than n structures");
        return nullptr;
    }
    napi_valuetype t;
    napi_typeof(env, a[4], &t);
    if (t != napi_null && !bytesOf(env, a[4], &fp, &fl)) return nullptr;
    napi_typeof(env, a[5], &t);
    if (t != napi_null && !bytesOf(env, a[5], &lp, &ll)) return nullptr;
    if ((fp && fl < (size_t)n * PT_OPT_NG * PT_OPT_NC * 8) || (lp && ll < (size_t)n * PT_OPT_LUX * PT_OPT_LUX * 8)) {
        napi_throw_range_error(env, nullptr, "far or lux smaller than n tallies");
        return nullptr;
    }
    double ms[2] = {0.0, 0.0};
    int rc = pt_optics_trace(dev, static_cast<const PtOpticsScene*>(sp), n, static_cast<PtOpticsResult*>(rp),
                             static_cast<double*>(fp), static_cast<double*>(lp), ms);
    if (rc) {
        std::string text = std::string("libpt: ") + pt_optics_error();
        napi_throw_error(env, nullptr, text.c_str());
        return nullptr;
    }
    napi_value arr;
    NAPI_OK(env, napi_create_array_with_length(env, 2, &arr));
    napi_set_element(env, arr, 0, num(env, ms[0]));
    napi_set_element(env, arr, 1, num(env, ms[1]));
    return arr;
}

napi_value developInfo(napi_env env, const PtDevelopInfo& di) {
    napi_value arr;
    NAPI_OK(env, napi_create_array_with_length(env, 4, &arr));
    napi_set_element(env, arr, 0, num(env, di.ev));
    napi_set_element(env, arr, 1, num(env, di.log2Average));
    napi_set_element(env, arr, 2, num(env, di.log2High));
    napi_set_element(env, arr, 3, num(env, di.metered));
    return arr;
}

// develop(h, settings (PtDevelop bytes), rgba (Uint8Array of w x h x 4)) -> [ev, log2Average, log2High, metered]
napi_value Develop(napi_env env, napi_callback_info info) {
    napi_value a[3];
    if (!args(env, info, 3, a)) return nullptr;
    pt_ctx* ctx = ctxOf(env, a[0]);
    if (!ctx) return nullptr;
    void *sp = nullptr, *op = nullptr;
    size_t sl = 0, ol = 0;
    if (!bytesOf(env, a[1], &sp, &sl) || !bytesOf(env, a[2], &op, &ol)) return nullptr;
    if (sl != sizeof(PtDevelop) || ol % 4) { napi_throw_range_error(env, nullptr, "settings not a PtDevelop, or rgba not 4 bytes a pixel"); return nullptr; }
    PtDevelopInfo di;
    int rc = pt_develop(ctx, static_cast<const PtDevelop*>(sp), sl, static_cast<uint8_t*>(op), ol / 4, &di);
    if (rc) return throwPt(env, ctx, rc);
    return developInfo(env, di);
}

// develop16(h, settings (PtDevelop bytes), rgba16 (Uint16Array of w x h x 4)) -> [ev, log2Average, log2High, metered]
napi_value Develop16(napi_env env, napi_callback_info info) {
    napi_value a[3];
    if (!args(env, info, 3, a)) return nullptr;
    pt_ctx* ctx = ctxOf(env, a[0]);
    if (!ctx) return nullptr;
    void *sp = nullptr, *op = nullptr;
    size_t sl = 0, ol = 0;
    if (!bytesOf(env, a[1], &sp, &sl) || !bytesOf(env, a[2], &op, &ol)) return nullptr;
    if (sl != sizeof(PtDevelop) || ol % 8) { napi_throw_range_error(env, nullptr, "settings not a PtDevelop, or rgba16 not 8 bytes a pixel"); return nullptr; }
    PtDevelopInfo di;
    int rc = pt_develop16(ctx, static_cast<const PtDevelop*>(sp), sl, static_cast<uint16_t*>(op), ol / 8, &di);
    if (rc) return throwPt(env, ctx, rc);
    return developInfo(env, di);
}

// developView(h, settings (PtDevelop bytes), view (Float32Array of 4), outW, outH, clipWarn, rgba (Uint8Array of
// outW x outH x 4)) -> [ev, log2Average, log2High, metered]
napi_value DevelopView(napi_env env, napi_callback_info info) {
    napi_value a[7];
    if (!args(env, info, 7, a)) return nullptr;
    pt_ctx* ctx = ctxOf(env, a[0]);
    if (!ctx) return nullptr;
    void *sp = nullptr, *vp = nullptr, *op = nullptr;
    size_t sl = 0, vl = 0, ol = 0;
    uint32_t w = 0, hh = 0, clip = 0;
    if (!bytesOf(env, a[1], &sp, &sl) || !bytesOf(env, a[2], &vp, &vl) || !bytesOf(env, a[6], &op, &ol)) return nullptr;
    if (napi_get_value_uint32(env, a[3], &w) != napi_ok || napi_get_value_uint32(env, a[4], &hh) != napi_ok ||
        napi_get_value_uint32(env, a[5], &clip) != napi_ok) {
        napi_throw_type_error(env, nullptr, "outW, outH, clipWarn: numbers");
        return nullptr;
    }
    if (sl != sizeof(PtDevelop) || vl != 16 || ol % 4) {
        napi_throw_range_error(env, nullptr, "settings not a PtDevelop, view not 4 floats, or rgba not 4 bytes a pixel");
        return nullptr;
    }
    PtDevelopInfo di;
    int rc = pt_develop_view(ctx, static_cast<const PtDevelop*>(sp), sl, static_cast<const float*>(vp), w, hh, clip,
                             static_cast<uint8_t*>(op), ol / 4, &di);
    if (rc) return throwPt(env, ctx, rc);
    return developInfo(env, di);
}

// developHistogram(h, settings (PtDevelop bytes), step, hist (Uint32Array of 4 x 256)) -> [ev, log2Average, log2High, metered]
napi_value DevelopHistogram(napi_env env, napi_callback_info info) {
    napi_value a[4];
    if (!args(env, info, 4, a)) return nullptr;
    pt_ctx* ctx = ctxOf(env, a[0]);
    if (!ctx) return nullptr;
    void *sp = nullptr, *hp = nullptr;
    size_t sl = 0, hl = 0;
    uint32_t step = 1;
    if (!bytesOf(env, a[1], &sp, &sl) || !bytesOf(env, a[3], &hp, &hl)) return nullptr;
    if (napi_get_value_uint32(env, a[2], &step) != napi_ok) { napi_throw_type_error(env, nullptr, "step: a number"); return nullptr; }
    if (sl != sizeof(PtDevelop) || hl != 4 * PT_HIST_BINS * 4) {
        napi_throw_range_error(env, nullptr, "settings not a PtDevelop, or hist not 1024 32-bit bins");
        return nullptr;
    }
    PtDevelopInfo di;
    int rc = pt_develop_histogram(ctx, static_cast<const PtDevelop*>(sp), sl, step, static_cast<uint32_t*>(hp), hl / 4, &di);
    if (rc) return throwPt(env, ctx, rc);
    return developInfo(env, di);
}

// opticsRun(h, sources, nsrc, trace, nfar, hitCap) -> Promise of { tally: ArrayBuffer (PtOpticsTally), far: Float64Array
// | null, hits: Float32Array }: the optics trace over the loaded scene on a worker thread (pt_optics_run). The inputs are
// copied in and the outputs made when it is done, so nothing of the caller's is held meanwhile.
struct OpticsRunWork {
    napi_async_work work;
    napi_deferred deferred;
    pt_ctx* ctx;
    std::vector<PtOpticsSource> src;
    PtOpticsTrace t;
    PtOpticsTally out;
    std::vector<double> far;
    std::vector<float> hits;
    uint32_t hitCap;
    int rc;
    std::string err;
};

void opticsRunExecute(napi_env, void* data) {
    OpticsRunWork* w = static_cast<OpticsRunWork*>(data);
    w->rc = pt_optics_run(w->ctx, w->src.data(), (uint32_t)w->src.size(), &w->t, sizeof w->t, &w->out, sizeof w->out,
                          w->far.empty() ? nullptr : w->far.data(), w->far.size(), w->hitCap ? w->hits.data() : nullptr, w->hitCap);
    if (w->rc) w->err = pt_last_error(w->ctx);
}

void opticsRunComplete(napi_env env, napi_status, void* data) {
    OpticsRunWork* w = static_cast<OpticsRunWork*>(data);
    if (w->rc) {
        napi_value msg, err;
        std::string text = "libpt: " + w->err;
        napi_create_string_utf8(env, text.c_str(), text.size(), &msg);
        napi_create_error(env, nullptr, msg, &err);
        napi_reject_deferred(env, w->deferred, err);
    } else {
        napi_value res, tally, far, hits, ab;
        void* d;
        napi_create_object(env, &res);
        napi_create_arraybuffer(env, sizeof w->out, &d, &tally);
        memcpy(d, &w->out, sizeof w->out);
        napi_set_named_property(env, res, "tally", tally);
        if (w->far.empty()) {
            napi_get_null(env, &far);
        } else {
            napi_create_arraybuffer(env, w->far.size() * 8, &d, &ab);
            memcpy(d, w->far.data(), w->far.size() * 8);
            napi_create_typedarray(env, napi_float64_array, w->far.size(), ab, 0, &far);
        }
        napi_set_named_property(env, res, "far", far);
        const size_t nh = (size_t)std::min<unsigned long long>(w->out.detectorHits, w->hitCap) * PT_OPT_HIT_WORDS;
        napi_create_arraybuffer(env, nh * 4 + 4, &d, &ab);
        if (nh) memcpy(d, w->hits.data(), nh * 4);
        napi_create_typedarray(env, napi_float32_array, nh, ab, 0, &hits);
        napi_set_named_property(env, res, "hits", hits);
        napi_resolve_deferred(env, w->deferred, res);
    }
    napi_delete_async_work(env, w->work);
    delete w;
}

napi_value OpticsRun(napi_env env, napi_callback_info info) {
    napi_value a[6];
    if (!args(env, info, 6, a)) return nullptr;
    pt_ctx* ctx = ctxOf(env, a[0]);
    if (!ctx) return nullptr;
    uint32_t nsrc = 0, nfar = 0, hitCap = 0;
    napi_get_value_uint32(env, a[2], &nsrc);
    napi_get_value_uint32(env, a[4], &nfar);
    napi_get_value_uint32(env, a[5], &hitCap);
    void *sp = nullptr, *tp = nullptr;
    size_t sl = 0, tl = 0;
    if (!bytesOf(env, a[1], &sp, &sl) || !bytesOf(env, a[3], &tp, &tl)) return nullptr;
    if (sl < (size_t)nsrc * sizeof(PtOpticsSource) || tl != sizeof(PtOpticsTrace)) {
        napi_throw_range_error(env, nullptr, "sources smaller than nsrc structures, or the trace not a PtOpticsTrace");
        return nullptr;
    }
    if (hitCap > (1u << 24)) { napi_throw_range_error(env, nullptr, "hitCap above 2^24"); return nullptr; }
    OpticsRunWork* w = new OpticsRunWork();
    w->ctx = ctx;
    w->src.assign(static_cast<const PtOpticsSource*>(sp), static_cast<const PtOpticsSource*>(sp) + nsrc);
    memcpy(&w->t, tp, sizeof w->t);
    w->far.assign(nfar, 0.0);
    w->hitCap = hitCap;
    w->hits.assign(size_t(hitCap) * PT_OPT_HIT_WORDS, 0.0f);
    w->rc = 0;
    napi_value promise, name;
    NAPI_OK(env, napi_create_promise(env, &w->deferred, &promise));
    napi_create_string_utf8(env, "pt_optics_run", NAPI_AUTO_LENGTH, &name);
    NAPI_OK(env, napi_create_async_work(env, nullptr, name, opticsRunExecute, opticsRunComplete, w, &w->work));
    NAPI_OK(env, napi_queue_async_work(env, w->work));
    return promise;
}

napi_value Init(napi_env env, napi_value exports) {
    napi_property_descriptor props[] = {
        {"create", nullptr, Create, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"destroy", nullptr, Destroy, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"abiVersion", nullptr, AbiVersion, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"structSize", nullptr, StructSize, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"compileScene", nullptr, CompileScene, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"setScene", nullptr, SetScene, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"setParams", nullptr, SetParams, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"foldParams", nullptr, FoldParams, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"setUniforms", nullptr, SetUniforms, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"clearAccum", nullptr, ClearAccum, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"trace", nullptr, Trace, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"opticsRun", nullptr, OpticsRun, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"develop", nullptr, Develop, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"developHistogram", nullptr, DevelopHistogram, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"develop16", nullptr, Develop16, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"developView", nullptr, DevelopView, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"readAccum", nullptr, ReadAccum, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"readStats", nullptr, ReadStats, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"opticsTrace", nullptr, OpticsTrace, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"autotune", nullptr, Autotune, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"setTuning", nullptr, SetTuning, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
    };
    napi_define_properties(env, exports, sizeof props / sizeof props[0], props);
    return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
