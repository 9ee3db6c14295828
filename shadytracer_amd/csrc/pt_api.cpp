// libpt: the host side of the C ABI (include/pt.h). It compiles each scene with hipRTC for gfx950 (device code
// embedded at build time from csrc/device/*.h, the emitted scene spliced in), caches the code objects, owns the
// structure-of-arrays path state in HBM and replays one captured hipGraph per sample.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <dlfcn.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pt.h"
#include "device/pt_dev.h"

namespace {

#include "pt_device_src.inc"   // kPtDevicePre, kPtDevicePost: the device sources as strings

const char* kKernelNames[] = {"pt_genrays", "pt_march", "pt_shade", "pt_occlude", "pt_resolve", "pt_mf_lut", "pt_mf_lut_avg", "pt_pick",
                              "pt_bin_count", "pt_bin_scan", "pt_bin_scatter", "pt_sppm_clear", "pt_photon_gen", "pt_surface",
                              "pt_guide_infer", "pt_guide_eval", "pt_guide_collect", "pt_guide_grad", "pt_guide_adam", "pt_guide_commit",
                              "pt_flush", "pt_opt_gen", "pt_opt_shade", "pt_opt_flush", "pt_ae_hist", "pt_ae_update",
                              "pt_develop", "pt_develop_hist", "pt_develop16",
                              "pt_develop_view"};
enum { K_GENRAYS, K_MARCH, K_SHADE, K_OCCLUDE, K_RESOLVE, K_MFLUT, K_MFAVG, K_PICK, K_BINCOUNT, K_BINSCAN, K_BINSCATTER,
       K_SPPMCLEAR, K_PHOTONGEN, K_SURFACE, K_GINFER, K_GEVAL, K_GCOLLECT, K_GGRAD, K_GADAM, K_GCOMMIT, K_FLUSH, K_OPTGEN, K_OPTSHADE,
       K_OPTFLUSH, K_AEHIST, K_AEUPDATE, K_DEVELOP, K_DEVHIST, K_DEVELOP16, K_DEVVIEW, K_COUNT };

// the guide network's sizes (device/pt_guide.h)
const unsigned int kGuideNP = 16169, kGuideWT = 17408, kGuideBias = 256, kGuideParts = 256, kGuideMaxSteps = 64;

thread_local std::string g_createError;

uint64_t fnv1a(const void* data, size_t n, uint64_t h = 1469598103934665603ull) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; ++i) { h ^= p[i]; h *= 1099511628211ull; }
    return h;
}

std::string libraryDir() {
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&fnv1a), &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        size_t s = p.rfind('/');
        if (s != std::string::npos) return p.substr(0, s);
    }
    return ".";
}

std::string defaultCacheDir(const char* given) {
    if (given && *given) return given;
    const char* env = getenv("PT_KCACHE");
    if (env && *env) return env;
    return libraryDir() + "/kcache";
}

void mkdirs(const std::string& d) {
    std::string cur;
    for (size_t i = 0; i < d.size(); ++i) {
        cur += d[i];
        if (d[i] == '/' && cur.size() > 1) mkdir(cur.c_str(), 0755);
    }
    mkdir(d.c_str(), 0755);
}

bool readFile(const std::string& path, std::vector<char>& out) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return false;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    out.resize(n > 0 ? size_t(n) : 0);
    size_t got = n > 0 ? fread(out.data(), 1, size_t(n), f) : 0;
    fclose(f);
    return got == out.size() && !out.empty();
}

bool writeFileAtomic(const std::string& path, const std::vector<char>& data) {
    std::string tmp = path + ".tmp" + std::to_string(getpid());
    FILE* f = fopen(tmp.c_str(), "wb");
    if (!f) return false;
    bool ok = fwrite(data.data(), 1, data.size(), f) == data.size();
    ok = (fclose(f) == 0) && ok;
    if (!ok) { unlink(tmp.c_str()); return false; }
    return rename(tmp.c_str(), path.c_str()) == 0;
}

// The atomic optimizer's DPP strategy scans a wave's divergent addends in a few DPP adds (the default loops over the
// active lanes), for the queue grabs whose first lane alone adds.
const char* kRtcOpts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                          "-mllvm", "-amdgpu-atomic-optimizer-strategy=DPP"};
const int kRtcNopts = sizeof(kRtcOpts) / sizeof(kRtcOpts[0]);

std::string fullSource(const char* scene) {
    std::string s;
    s.reserve(strlen(kPtDevicePre) + strlen(scene) + strlen(kPtDevicePost) + 64);
    s += kPtDevicePre;
    s += "\n// ---- emitted scene ----\n";
    s += scene;
    s += "\n// ---- end of scene ----\n";
    s += kPtDevicePost;
    return s;
}

// Compiles (or finds) the code object for a scene. Returns false with the log in err.
bool compileScene(const char* scene, const std::string& cacheDir, std::vector<char>& code, std::string& path, std::string& err) {
    std::string src = fullSource(scene);
    uint64_t h = fnv1a(src.data(), src.size());
    for (int i = 0; i < kRtcNopts; ++i) h = fnv1a(kRtcOpts[i], strlen(kRtcOpts[i]), h);
    int ver_major = 0, ver_minor = 0;
    hiprtcVersion(&ver_major, &ver_minor);
    h = fnv1a(&ver_major, sizeof ver_major, h);
    h = fnv1a(&ver_minor, sizeof ver_minor, h);
    char name[64];
    snprintf(name, sizeof name, "pt-%016llx.hsaco", (unsigned long long)h);
    path = cacheDir + "/" + name;
    if (readFile(path, code)) return true;
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "pt_scene.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        err = "hiprtcCreateProgram failed";
        return false;
    }
    hiprtcResult rc = hiprtcCompileProgram(prog, kRtcNopts, kRtcOpts);
    size_t logSize = 0;
    hiprtcGetProgramLogSize(prog, &logSize);
    std::string log(logSize, '\0');
    if (logSize) hiprtcGetProgramLog(prog, &log[0]);
    if (rc != HIPRTC_SUCCESS) {
        err = std::string("scene compile failed: ") + hiprtcGetErrorString(rc) + "\n" + log;
        hiprtcDestroyProgram(&prog);
        return false;
    }
    size_t codeSize = 0;
    hiprtcGetCodeSize(prog, &codeSize);
    code.resize(codeSize);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    mkdirs(cacheDir);
    writeFileAtomic(path, code);   // a cache miss that cannot be stored still loads
    return true;
}

}  // namespace

struct pt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::string cacheDir;
    unsigned int cus = 256, blocksPerCu = 8;
    unsigned int stragglers = PT_STRAGGLER_ROUNDS;   // rounds beyond maxbounce + 1 (pt_opts.straggler_rounds)
    bool direct = false;      // PT_OPT_DIRECT_LAUNCH: samples launched kernel by kernel, no graph replay
    PtTuning tune{};          // the persistent kernels' launch choices (pt_set_tuning, pt_autotune)

    hipModule_t module = nullptr;
    hipFunction_t fn[K_COUNT] = {};
    bool building = false;
    std::string sceneSrc;            // the loaded scene's emitted source and parameter block, for pt_fold_params
    std::vector<float> params;
    hipModule_t foldedModule = nullptr;   // the scene compiled with those parameters as constants, while in use
    hipFunction_t genericFn[K_COUNT] = {};

    PtDev host{};           // host mirror of the pointer table
    PtDev* dev = nullptr;   // device copy
    bool devDirty = true;

    unsigned int capacity = 0;   // path slots allocated
    unsigned int pixels = 0;     // pixels the frame/accumulators hold
    unsigned int nparams = 0;

    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    uint64_t graphKey = 0;       // what the captured graph's launches were sized for (rounds, rays, photons)

    unsigned int sppmCapacity = 0;   // photon records allocated

    // the developed picture (pt_develop): the meter's histogram and state, the settings and the 8-bit picture on the device
    unsigned int* aeHist = nullptr;
    float* aeState = nullptr;
    PtDevelop* devSet = nullptr;
    unsigned char* devRgba = nullptr;
    size_t devRgbaPx = 0;

    // the optics trace (pt_optics_run): its table on the device and the buffers it points at
    PtOptDev opt{};
    PtOptDev* optDev = nullptr;
    unsigned int optSrcCap = 0, optFarCap = 0, optHitCap = 0;

    bool guideReady = false;          // guide network allocated and initialised
    unsigned int guideMixSlots = 0;   // slots gMix holds
    unsigned int guideRecPaths = 0, guideRecVerts = 0, guideBatchCap = 0;

    std::vector<void*> owned;
};

namespace {

int fail(pt_ctx* c, int code, const char* fmt, ...) {
    char buf[4096];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf; else g_createError = buf;
    return code;
}

#define HIPCHK(c, expr)                                                                             \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return fail((c), PT_ERR_HIP, "%s: %s (%s:%d)", #expr,                 \
                                          hipGetErrorString(e_), __FILE__, __LINE__);               \
    } while (0)

void freeBuf(pt_ctx* c, void*& p) {
    if (!p) return;
    hipFree(p);
    for (auto& q : c->owned) if (q == p) q = nullptr;
    p = nullptr;
}

template <typename T>
int allocBuf(pt_ctx* c, T*& p, size_t bytes) {
    void* raw = nullptr;
    hipError_t e = hipMalloc(&raw, bytes ? bytes : 16);
    if (e != hipSuccess) return fail(c, PT_ERR_HIP, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    p = static_cast<T*>(raw);
    c->owned.push_back(raw);
    return PT_OK;
}

void destroyGraph(pt_ctx* c) {
    if (c->graphExec) { hipGraphExecDestroy(c->graphExec); c->graphExec = nullptr; }
    if (c->graph) { hipGraphDestroy(c->graph); c->graph = nullptr; }
}

// The reference's raygen tile grid (core/genrays.js:22-34): rounded, so a 1-spp render is one tile per pixel.
void raygenTiles(unsigned int N, unsigned int w, unsigned int h, unsigned int& tx, unsigned int& ty) {
    double t = std::round(std::sqrt(double(N) * (double(w) / double(h ? h : 1))));
    tx = t < 1 ? 1u : (unsigned int)t;
    ty = (N + tx - 1) / tx;
    if (ty < 1) ty = 1;
}

int ensureSlots(pt_ctx* c, unsigned int slots);

// Frame and accumulators for w x h pixels; path slots for the larger of the pixels and the photons a pass emits; the
// photon map (three tables of 2^20 words; room for two records, 32 bytes and two links each, a photon) when photons
// are asked for.
int ensureBuffers(pt_ctx* c, unsigned int w, unsigned int h, unsigned int photons) {
    unsigned int px = w * h;
    if (px == 0) return fail(c, PT_ERR_ARG, "width and height must be positive (got %u x %u)", w, h);
    if (px > c->pixels) {
        destroyGraph(c);
        void* p;
        p = c->host.frame; freeBuf(c, p);
        p = c->host.accum; freeBuf(c, p);
        p = c->host.moments; freeBuf(c, p);
        int rc;
        if ((rc = allocBuf(c, c->host.frame, size_t(px) * 16))) return rc;
        if ((rc = allocBuf(c, c->host.accum, size_t(px) * 32))) return rc;
        if ((rc = allocBuf(c, c->host.moments, size_t(px) * 16))) return rc;
        HIPCHK(c, hipMemsetAsync(c->host.frame, 0, size_t(px) * 16, c->stream));
        HIPCHK(c, hipMemsetAsync(c->host.accum, 0, size_t(px) * 32, c->stream));
        HIPCHK(c, hipMemsetAsync(c->host.moments, 0, size_t(px) * 16, c->stream));
        c->pixels = px;
    }
    unsigned int slots = photons > px ? photons : px;
    // a photon stores once at each evaluable surface it meets along its delta chain: room for two a photon
    unsigned int records = photons > (1u << 30) ? (1u << 31) : 2u * photons;
    if (records > c->sppmCapacity) {
        destroyGraph(c);
        void* p;
        p = c->host.sppmRecs; freeBuf(c, p); c->host.sppmRecs = nullptr;
        p = c->host.sppmNext; freeBuf(c, p); c->host.sppmNext = nullptr;
        int rc;
        // heads: the cells' lists, the cells' counts, the sub-lists; a record's links: its list's next, its sub-list's
        if (!c->host.sppmHeads && (rc = allocBuf(c, c->host.sppmHeads, size_t(3) * PT_SPPM_HEADS * 4))) return rc;
        if ((rc = allocBuf(c, c->host.sppmRecs, size_t(records) * 32))) return rc;
        if ((rc = allocBuf(c, c->host.sppmNext, size_t(records) * 8))) return rc;
        c->sppmCapacity = records;
    }
    c->host.sppmCap = c->sppmCapacity;
    return ensureSlots(c, slots);
}

// Path slots (state, queues, masks, hit records, shadow rays) for `slots` rays at once; grows, never shrinks.
int ensureSlots(pt_ctx* c, unsigned int slots) {
    if (slots > c->capacity) {
        destroyGraph(c);
        float** f4[] = {&c->host.pos, &c->host.dir, &c->host.beta, &c->host.rad, &c->host.aux, &c->host.shOrig, &c->host.shDir, &c->host.shRad,
                        &c->host.hitRec};
        for (auto pp : f4) { void* p = *pp; freeBuf(c, p); *pp = nullptr; }
        unsigned int** u1[] = {&c->host.qMarch0, &c->host.qMarch1, &c->host.qShadow, &c->host.qSorted};
        for (auto pp : u1) { void* p = *pp; freeBuf(c, p); *pp = nullptr; }
        { void* p = c->host.meta; freeBuf(c, p); c->host.meta = nullptr; }
        { void* p = c->host.maskBuf; freeBuf(c, p); c->host.maskBuf = nullptr; }
        int rc;
        for (auto pp : f4) if ((rc = allocBuf(c, *pp, size_t(slots) * 16))) return rc;
        if ((rc = allocBuf(c, c->host.meta, size_t(slots) * 16))) return rc;
        if ((rc = allocBuf(c, c->host.maskBuf, size_t(slots) * 8))) return rc;
        for (auto pp : u1) if ((rc = allocBuf(c, *pp, size_t(slots) * 4))) return rc;
        c->capacity = slots;
    }
    c->host.capacity = c->capacity;
    c->devDirty = true;
    return PT_OK;
}

int syncDev(pt_ctx* c) {
    if (!c->devDirty) return PT_OK;
    HIPCHK(c, hipMemcpyAsync(c->dev, &c->host, sizeof(PtDev), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->devDirty = false;
    return PT_OK;
}

// maxbounce + 1 rounds end every path that bounces each round; the straggler rounds end those that met hits from
// behind (moved on without a bounce), and the flush closes what is left
unsigned int rounds(const pt_ctx* c) { return c->host.u.maxbounce + 1u + c->stragglers; }

// the refill thresholds the kernels read: the uniforms' marchRefill where set, else the tuning's
void applyRefills(pt_ctx* c) {
    const unsigned int r = c->host.u.marchRefill;
    c->host.refillMarch = r ? r : c->tune.refillMarch;
    c->host.refillOcclude = r ? r : c->tune.refillOcclude;
    c->devDirty = true;
}

hipError_t launch(pt_ctx* c, int k, unsigned int grid, unsigned int qsel) {
    void* args1[] = {&c->dev};
    void* args2[] = {&c->dev, &qsel};
    bool two = (k == K_MARCH || k == K_SHADE || k == K_BINCOUNT || k == K_BINSCATTER || k == K_OCCLUDE || k == K_SURFACE ||
                k == K_GINFER || k == K_GGRAD || k == K_GADAM || k == K_GCOMMIT || k == K_FLUSH);
    unsigned int block = k == K_BINSCAN ? PT_BINS : k == K_GINFER ? 512u : (k == K_GGRAD || k == K_GCOMMIT) ? 64u : 256u;
    return hipModuleLaunchKernel(c->fn[k], grid, 1, 1, block, 1, 1, 0, c->stream, two ? args2 : args1, nullptr);
}

unsigned int gridFor(const pt_ctx* c, int k) {
    switch (k) {
        case K_GENRAYS: return (c->host.nrays + 255u) / 256u;
        case K_RESOLVE: return (c->host.u.width * c->host.u.height + 255u) / 256u;
        case K_BINCOUNT: case K_BINSCATTER: return PT_BIN_BLOCKS;
        case K_BINSCAN: return 1;
        case K_SPPMCLEAR: return 3u * PT_SPPM_HEADS / 256u;
        case K_PHOTONGEN: return (c->host.nphotons + 255u) / 256u;
        case K_GINFER: return c->cus;
        case K_GCOLLECT: return (c->host.gPaths * c->host.gMaxV + 255u) / 256u;
        case K_GGRAD: return kGuideParts;
        case K_GADAM: return (kGuideNP + 255u) / 256u;
        case K_GCOMMIT: return 1u;
        default: return c->host.gridBlocks;
    }
}

// One sample's launches, in order; `events` (2 per launch) brackets each when given.
hipError_t enqueueSample(pt_ctx* c, std::vector<hipEvent_t>* events, std::vector<int>* kinds) {
    hipError_t e;
    auto one = [&](int k, unsigned int qsel) -> hipError_t {
        if (events && kinds->size() * 2 + 1 >= events->size()) return hipErrorInvalidValue;   // more launches than timed
        if (events) { hipEventRecord((*events)[kinds->size() * 2], c->stream); }
        hipError_t r = launch(c, k, gridFor(c, k), qsel);
        if (events) { hipEventRecord((*events)[kinds->size() * 2 + 1], c->stream); kinds->push_back(k); }
        return r;
    };
    unsigned int R = rounds(c);
    const bool guiding = c->host.u.guide != 0u && c->guideReady;
    if (c->host.nphotons) {
        // THE PHOTON PASS: the map cleared and r_k set, photons emitted, marched and shaded round by round. Without
        // light tracing there are no shadow rays, and occlude runs on an empty queue for its resets of the march queue;
        // with it the photons' connections to the camera are the shadow queue, binned as the camera's are.
        const bool lt = c->host.u.lt != 0u;
        if ((e = one(K_SPPMCLEAR, 0)) || (e = one(K_PHOTONGEN, 0))) return e;
        for (unsigned int it = 0; it < R; ++it) {
            unsigned int in = it & 1u;
            if ((e = one(K_BINCOUNT, in)) || (e = one(K_BINSCAN, 0)) || (e = one(K_BINSCATTER, in))) return e;
            if ((e = one(K_MARCH, in)) || (e = one(K_SURFACE, in ^ 1u)) || (e = one(K_SHADE, in ^ 1u))) return e;
            if (lt && ((e = one(K_BINCOUNT, 2)) || (e = one(K_BINSCAN, 0)) || (e = one(K_BINSCATTER, 2)))) return e;
            if ((e = one(K_OCCLUDE, in))) return e;
        }
    }
    if ((e = one(K_GENRAYS, 0))) return e;
    for (unsigned int it = 0; it < R; ++it) {
        unsigned int in = it & 1u;
        if ((e = one(K_BINCOUNT, in)) || (e = one(K_BINSCAN, 0)) || (e = one(K_BINSCATTER, in))) return e;
        if ((e = one(K_MARCH, in)) || (e = one(K_SURFACE, in ^ 1u))) return e;
        if (guiding && (e = one(K_GINFER, in ^ 1u))) return e;
        if ((e = one(K_SHADE, in ^ 1u))) return e;
        if ((e = one(K_BINCOUNT, 2)) || (e = one(K_BINSCAN, 0)) || (e = one(K_BINSCATTER, 2))) return e;
        if ((e = one(K_OCCLUDE, in))) return e;
    }
    // the last round's shade filled march queue R & 1 with the paths still going
    if ((e = one(K_FLUSH, R & 1u)) || (e = one(K_RESOLVE, 0))) return e;
    if (guiding) {
        // THE GUIDE'S TRAINING: the sample's recorded vertices into samples, then its Adam steps
        if ((e = one(K_GCOLLECT, 0))) return e;
        for (unsigned int j = 0; j < c->host.g.steps; ++j)
            if ((e = one(K_GGRAD, j)) || (e = one(K_GADAM, j)) || (e = one(K_GCOMMIT, j))) return e;
    }
    return hipSuccess;
}

uint64_t graphKey(const pt_ctx* c) {
    // what the captured launches were sized for
    const unsigned int k[] = {rounds(c), c->host.nrays, c->host.nphotons, c->host.gridBlocks,
                              (c->host.u.guide && c->guideReady) ? 1u : 0u, c->host.g.steps, c->host.gPaths, c->host.gMaxV,
                              c->host.u.lt ? 1u : 0u};
    return fnv1a(k, sizeof k);
}

int buildGraph(pt_ctx* c) {
    if (c->direct) return PT_OK;                 // launched kernel by kernel (launchSample)
    if (c->graphExec && c->graphKey == graphKey(c)) return PT_OK;
    destroyGraph(c);
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = enqueueSample(c, nullptr, nullptr);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return fail(c, PT_ERR_HIP, "graph capture failed: %s / %s", hipGetErrorString(e), hipGetErrorString(e2));
    }
    c->graph = g;
    HIPCHK(c, hipGraphInstantiate(&c->graphExec, g, nullptr, nullptr, 0));
    c->graphKey = graphKey(c);
    return PT_OK;
}

// One sample on the stream: the captured graph replayed, or with PT_OPT_DIRECT_LAUNCH its launches one by one
hipError_t launchSample(pt_ctx* c) {
    return c->direct ? enqueueSample(c, nullptr, nullptr) : hipGraphLaunch(c->graphExec, c->stream);
}

// ---- the guide network's host side
uint16_t bf16Bits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return 0x7FC0u;
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
    return (uint16_t)(u >> 16);
}

// He-uniform weights from a PCG32 stream of the seed (biases 0; the head's layer at a tenth, so the first mixture is
// broad), the engine's parameter order.
void guideInit(uint32_t seed, std::vector<float>& th) {
    th.assign(kGuideNP, 0.0f);
    uint64_t state = 0x853C49E6748FEA9Bull ^ (uint64_t(seed) << 1), inc = (uint64_t(seed) << 1) | 1u;
    auto next = [&]() {
        uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        uint32_t x = (uint32_t)(((old >> 18u) ^ old) >> 27u), r = (uint32_t)(old >> 59u);
        return ((x >> r) | (x << ((32u - r) & 31u))) * (1.0 / 4294967296.0);
    };
    const unsigned int sizes[5] = {80, 64, 64, 64, 41};
    size_t o = 0;
    for (int L = 0; L < 4; ++L) {
        const double lim = std::sqrt(6.0 / sizes[L]) * (L == 3 ? 0.1 : 1.0);
        for (unsigned int k = 0; k < sizes[L] * sizes[L + 1]; ++k) th[o++] = (float)((2.0 * next() - 1.0) * lim);
        o += sizes[L + 1];
    }
}

// Master weights in: Adam's state cleared, the BF16 transposed copy and the biases written, the step count zeroed.
int guideUpload(pt_ctx* c, const float* th) {
    std::vector<uint16_t> w(kGuideWT, 0);
    std::vector<float> b(kGuideBias, 0.0f);
    const unsigned int sizes[5] = {80, 64, 64, 64, 41}, wt[4] = {0, 5120, 9216, 13312};
    size_t o = 0;
    for (int L = 0; L < 4; ++L) {
        const unsigned int in = sizes[L], out = sizes[L + 1];
        for (unsigned int i = 0; i < in; ++i)
            for (unsigned int j = 0; j < out; ++j) w[wt[L] + j * in + i] = bf16Bits(th[o + i * out + j]);
        o += in * out;
        for (unsigned int j = 0; j < out; ++j) b[64 * L + j] = th[o + j];
        o += out;
    }
    HIPCHK(c, hipMemcpyAsync(c->host.gMaster, th, kGuideNP * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(c->host.gM, 0, kGuideNP * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(c->host.gV, 0, kGuideNP * 4, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->host.gW, w.data(), kGuideWT * 2, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->host.gBias, b.data(), kGuideBias * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(c->host.counters + PT_C_GUIDE_T, 0, 4, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PT_OK;
}

// The guide's buffers for the current image and settings: the network's (once), the mixtures by slot, the recorded
// paths' vertices, the training samples.
int ensureGuide(pt_ctx* c) {
    int rc;
    PtDev& h = c->host;
    if (!h.gMaster) {
        if ((rc = allocBuf(c, h.gMaster, kGuideNP * 4)) || (rc = allocBuf(c, h.gM, kGuideNP * 4)) ||
            (rc = allocBuf(c, h.gV, kGuideNP * 4)) || (rc = allocBuf(c, h.gPart, size_t(kGuideParts) * kGuideNP * 4)) ||
            (rc = allocBuf(c, h.gW, kGuideWT * 2)) || (rc = allocBuf(c, h.gBias, kGuideBias * 4)) ||
            (rc = allocBuf(c, h.gSumF, kGuideMaxSteps * 8)) || (rc = allocBuf(c, h.gLoss, kGuideMaxSteps * 8)))
            return rc;
        HIPCHK(c, hipMemsetAsync(h.gSumF, 0, kGuideMaxSteps * 8, c->stream));
        HIPCHK(c, hipMemsetAsync(h.gLoss, 0, kGuideMaxSteps * 8, c->stream));
        std::vector<float> th;
        guideInit(h.g.seed, th);
        if ((rc = guideUpload(c, th.data()))) return rc;
    }
    if (c->capacity > c->guideMixSlots) {
        destroyGraph(c);
        void* p = h.gMix; freeBuf(c, p); h.gMix = nullptr;
        if ((rc = allocBuf(c, h.gMix, size_t(c->capacity) * 176))) return rc;
        c->guideMixSlots = c->capacity;
    }
    // which paths record: at about 4 training samples a recorded path, batch x steps of them a sample (each step draws
    // its batch from all of them); room for four times that
    const unsigned long long want = (unsigned long long)h.g.batch * h.g.steps;
    unsigned int stride = h.g.stride ? h.g.stride : (unsigned int)std::max<unsigned long long>(1ull, (4ull * h.nrays) / std::max(want, 1ull));
    const unsigned int paths = (h.nrays + stride - 1u) / stride, verts = h.u.maxbounce + 1u;
    if (paths > c->guideRecPaths || verts > c->guideRecVerts) {
        destroyGraph(c);
        void* p;
        p = h.gRec; freeBuf(c, p); h.gRec = nullptr;
        p = h.gPath; freeBuf(c, p); h.gPath = nullptr;
        p = h.gVerts; freeBuf(c, p); h.gVerts = nullptr;
        const unsigned int np = std::max(paths, c->guideRecPaths), nv = std::max(verts, c->guideRecVerts);
        if ((rc = allocBuf(c, h.gRec, size_t(np) * nv * 7 * 16)) || (rc = allocBuf(c, h.gPath, size_t(np) * 32)) ||
            (rc = allocBuf(c, h.gVerts, size_t(np) * 4)))
            return rc;
        HIPCHK(c, hipMemsetAsync(h.gVerts, 0, size_t(np) * 4, c->stream));
        c->guideRecPaths = np;
        c->guideRecVerts = nv;
    }
    h.gStride = stride;
    h.gPaths = paths;
    h.gMaxV = verts;
    const unsigned int cap = (unsigned int)std::min<unsigned long long>(4ull * want, 1ull << 26);
    if (cap > c->guideBatchCap) {
        void* p = h.gBatch; freeBuf(c, p); h.gBatch = nullptr;
        if ((rc = allocBuf(c, h.gBatch, size_t(cap) * 64))) return rc;
        c->guideBatchCap = cap;
    }
    h.gBatchCap = cap;
    c->guideReady = true;
    c->devDirty = true;
    return PT_OK;
}

void setGuideDefaults(PtGuideConfig& g) {
    memset(&g, 0, sizeof g);
    for (int a = 0; a < 3; ++a) { g.boxMin[a] = -10.0f; g.boxMax[a] = 10.0f; }
    g.lr = 1e-2f;
    g.batch = 1u << 16;
    g.steps = 4;
    g.stride = 0;
    g.warmup = 16;
    g.seed = 1;
}

void setDefaults(PtUniforms& u) {
    memset(&u, 0, sizeof u);
    u.width = 256; u.height = 256;
    u.maxbounce = 10; u.maxiter = 400; u.stepscale = 0.9f;
    u.eps = 0.0f; u.epsAuto = 1e-4f; u.envRadius = 400.0f;
    u.nee = 1; u.rr = 0; u.rrStart = 3; u.diffuseCap = 0;
    u.fireflyClamp = 0.0f; u.fireflyWhite = 1.0f; u.maxaverages = 0; u.seed = 20261203u;
    u.cam.pos[1] = -3.0f; u.cam.pos[2] = 1.5f;
    u.cam.right[0] = 1.0f; u.cam.up[2] = 1.0f; u.cam.fwd[1] = -1.0f;
    u.cam.fx = 0.57735f; u.cam.fy = 0.57735f;
    u.cam.focus = 1.0f;
    u.faceShare = 0.5f;
    u.marchRefill = 0;
    u.guideFrac = 0.9f;
    u.sppmRadius = 0.02f;
    u.sppmAlpha = 2.0f / 3.0f;
}

}  // namespace

extern "C" {

int pt_abi_version(void) { return PT_ABI_VERSION; }

size_t pt_struct_size(int which) {
    switch (which) {
        case 0: return sizeof(PtUniforms);
        case 1: return sizeof(PtMaterial);
        case 2: return sizeof(PtLightFace);
        case 3: return sizeof(PtCamera);
        case 4: return sizeof(PtEnv);
        case 5: return sizeof(PtHaze);
        case 6: return sizeof(pt_scene_desc);
        case 7: return sizeof(pt_opts);
        case 8: return sizeof(PtGuideConfig);
        case 9: return sizeof(PtOpticsScene);
        case 10: return sizeof(PtOpticsResult);
        case 11: return sizeof(PtTuning);
        case 12: return sizeof(PtOpticsSource);
        case 13: return sizeof(PtOpticsTrace);
        case 14: return sizeof(PtOpticsTally);
        case 15: return sizeof(PtDevelop);
        case 16: return sizeof(PtDevelopInfo);
        default: return 0;
    }
}

const char* pt_last_error(const pt_ctx* ctx) { return ctx ? ctx->err.c_str() : g_createError.c_str(); }

int pt_compile_scene(const char* hip_src, const char* cache_dir, char* path_out, size_t path_cap) {
    if (!hip_src) return fail(nullptr, PT_ERR_ARG, "hip_src is NULL");
    std::vector<char> code;
    std::string path, err;
    if (!compileScene(hip_src, defaultCacheDir(cache_dir), code, path, err)) return fail(nullptr, PT_ERR_COMPILE, "%s", err.c_str());
    if (path_out && path_cap) snprintf(path_out, path_cap, "%s", path.c_str());
    return PT_OK;
}

int pt_create(pt_ctx** out, const pt_opts* opts) {
    if (!out) return fail(nullptr, PT_ERR_ARG, "out is NULL");
    *out = nullptr;
    pt_ctx* c = new pt_ctx();
    c->device = opts ? opts->device : 0;
    c->cacheDir = defaultCacheDir(opts ? opts->cache_dir : nullptr);
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= c->device) {
        int rc = fail(nullptr, PT_ERR_HIP, "no HIP device %d (%d found): %s", c->device, ndev, hipGetErrorString(e));
        delete c;
        return rc;
    }
    if ((e = hipSetDevice(c->device)) != hipSuccess || (e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) {
        int rc = fail(nullptr, PT_ERR_HIP, "device setup: %s", hipGetErrorString(e));
        delete c;
        return rc;
    }
    int cus = 0;
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device);
    c->cus = cus > 0 ? (unsigned int)cus : 256u;
    c->blocksPerCu = (opts && opts->blocks_per_cu) ? opts->blocks_per_cu : 8u;
    c->stragglers = !(opts && opts->straggler_rounds) ? PT_STRAGGLER_ROUNDS
                  : opts->straggler_rounds == PT_STRAGGLER_NONE ? 0u : std::min(opts->straggler_rounds, 16u);
    c->direct = opts && (opts->flags & PT_OPT_DIRECT_LAUNCH);
    setDefaults(c->host.u);
    setGuideDefaults(c->host.g);
    c->host.gridBlocks = c->cus * c->blocksPerCu;
    c->tune.blocksPerCu = c->blocksPerCu;
    c->tune.refillMarch = c->tune.refillOcclude = 16u;
    applyRefills(c);
    int rc;
    if ((rc = allocBuf(c, c->dev, sizeof(PtDev))) || (rc = allocBuf(c, c->host.counters, PT_C_WORDS * 4)) ||
        (rc = allocBuf(c, c->host.stats, PT_S_WORDS * 8)) || (rc = allocBuf(c, c->host.mf, PT_MF_WORDS * 4)) ||
        (rc = allocBuf(c, c->host.mats, PT_MAX_MATERIALS * sizeof(PtMaterial))) ||
        (rc = allocBuf(c, c->host.faces, PT_MAX_FACES * sizeof(PtLightFace))) ||
        (rc = allocBuf(c, c->host.binHist, PT_BIN_BLOCKS * PT_BINS * 4)) || (rc = allocBuf(c, c->host.binOff, PT_BIN_BLOCKS * PT_BINS * 4))) {
        g_createError = c->err;
        pt_destroy(c);
        return rc;
    }
    hipMemsetAsync(c->host.counters, 0, PT_C_WORDS * 4, c->stream);
    hipMemsetAsync(c->host.stats, 0, PT_S_WORDS * 8, c->stream);
    if ((rc = ensureBuffers(c, c->host.u.width, c->host.u.height, 0))) {
        g_createError = c->err;
        pt_destroy(c);
        return rc;
    }
    *out = c;
    return PT_OK;
}

void pt_destroy(pt_ctx* c) {
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    destroyGraph(c);
    for (void* p : c->owned) if (p) hipFree(p);
    if (c->foldedModule) hipModuleUnload(c->foldedModule);
    if (c->module) hipModuleUnload(c->module);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

// Compiles (or finds) and loads a scene's code object and its kernels; on failure nothing is kept.
static int loadSceneModule(pt_ctx* c, const char* src, hipModule_t& mod, hipFunction_t (&fns)[K_COUNT]) {
    std::vector<char> code;
    std::string path, err;
    if (!compileScene(src, c->cacheDir, code, path, err)) return fail(c, PT_ERR_COMPILE, "%s", err.c_str());
    mod = nullptr;
    hipError_t e = hipModuleLoadData(&mod, code.data());
    if (e != hipSuccess) return fail(c, PT_ERR_COMPILE, "hipModuleLoadData(%s): %s", path.c_str(), hipGetErrorString(e));
    for (int k = 0; k < K_COUNT; ++k) {
        if ((e = hipModuleGetFunction(&fns[k], mod, kKernelNames[k])) != hipSuccess) {
            hipModuleUnload(mod);
            mod = nullptr;
            return fail(c, PT_ERR_COMPILE, "kernel %s missing: %s", kKernelNames[k], hipGetErrorString(e));
        }
    }
    return PT_OK;
}

// Back to the kernels that read the parameter block (after the block changes, or a new scene).
static void unfold(pt_ctx* c) {
    if (!c->foldedModule) return;
    hipStreamSynchronize(c->stream);
    destroyGraph(c);
    memcpy(c->fn, c->genericFn, sizeof c->fn);
    hipModuleUnload(c->foldedModule);
    c->foldedModule = nullptr;
}

int pt_set_scene(pt_ctx* c, const pt_scene_desc* s) {
    if (!c || !s || !s->hip_src) return fail(c, PT_ERR_ARG, "scene or its hip_src is NULL");
    if (s->nmaterials == 0 || s->nmaterials > PT_MAX_MATERIALS) return fail(c, PT_ERR_ARG, "materials: %u (1..%d)", s->nmaterials, PT_MAX_MATERIALS);
    if (s->nfaces > PT_MAX_FACES) return fail(c, PT_ERR_ARG, "light faces: %u (max %d)", s->nfaces, PT_MAX_FACES);
    if (s->nparams && !s->params) return fail(c, PT_ERR_ARG, "params is NULL");
    for (uint32_t i = 0; i < s->nmaterials; ++i)
        if (s->materials[i].face >= (int)s->nfaces) return fail(c, PT_ERR_ARG, "material %u names light face %d of %u", i, s->materials[i].face, s->nfaces);
    c->building = true;
    hipModule_t mod = nullptr;
    hipFunction_t fns[K_COUNT];
    int lrc = loadSceneModule(c, s->hip_src, mod, fns);
    c->building = false;
    if (lrc) return lrc;
    unfold(c);
    hipStreamSynchronize(c->stream);
    destroyGraph(c);
    if (c->module) hipModuleUnload(c->module);
    c->module = mod;
    memcpy(c->fn, fns, sizeof fns);
    memcpy(c->genericFn, fns, sizeof fns);
    c->sceneSrc = s->hip_src;
    c->params.assign(s->params, s->params + s->nparams);
    // tables
    if (s->nparams > c->nparams) {
        void* p = (void*)c->host.params; freeBuf(c, p); c->host.params = nullptr;
        float* np = nullptr;
        int rc;
        if ((rc = allocBuf(c, np, size_t(s->nparams) * 4))) return rc;
        c->host.params = np;
    }
    c->nparams = s->nparams;
    if (!c->host.params) { float* np = nullptr; int rc; if ((rc = allocBuf(c, np, 16))) return rc; c->host.params = np; }
    if (s->nparams) HIPCHK(c, hipMemcpyAsync((void*)c->host.params, s->params, size_t(s->nparams) * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync((void*)c->host.mats, s->materials, size_t(s->nmaterials) * sizeof(PtMaterial), hipMemcpyHostToDevice, c->stream));
    if (s->nfaces) HIPCHK(c, hipMemcpyAsync((void*)c->host.faces, s->faces, size_t(s->nfaces) * sizeof(PtLightFace), hipMemcpyHostToDevice, c->stream));
    c->host.nmats = s->nmaterials;
    c->host.nfaces = s->nfaces;
    float tot = 0.0f;
    for (uint32_t i = 0; i < s->nfaces; ++i) tot += s->faces[i].weight > 0.0f ? s->faces[i].weight : 0.0f;
    c->host.faceTotal = tot;
    c->devDirty = true;
    if (!c->host.mfReady) {
        void* args[] = {&c->host.mf};
        unsigned int n = PT_MF_NIOR * PT_MF_NR * PT_MF_NMU + PT_MF_NR * PT_MF_NMU;
        HIPCHK(c, hipModuleLaunchKernel(c->fn[K_MFLUT], (n + 255) / 256, 1, 1, 256, 1, 1, 0, c->stream, args, nullptr));
        HIPCHK(c, hipModuleLaunchKernel(c->fn[K_MFAVG], (PT_MF_NIOR * PT_MF_NR + 255) / 256, 1, 1, 256, 1, 1, 0, c->stream, args, nullptr));
        c->host.mfReady = 1;
    }
    if (s->flags & PT_SCENE_KEEP_PICTURE) return PT_OK;
    return pt_clear_accum(c);
}

int pt_set_params(pt_ctx* c, const float* params, uint32_t n) {
    if (!c || (n && !params)) return fail(c, PT_ERR_ARG, "params is NULL");
    if (!c->module) return fail(c, PT_ERR_STATE, "no scene loaded");
    if (n != c->nparams) return fail(c, PT_ERR_ARG, "params: %u values, the scene has %u", n, c->nparams);
    unfold(c);
    if (n) HIPCHK(c, hipMemcpyAsync((void*)c->host.params, params, size_t(n) * 4, hipMemcpyHostToDevice, c->stream));
    c->params.assign(params, params + n);
    return PT_OK;
}

int pt_fold_params(pt_ctx* c) {
    if (!c) return fail(c, PT_ERR_ARG, "ctx is NULL");
    if (!c->module) return fail(c, PT_ERR_STATE, "no scene loaded");
    if (c->foldedModule) return PT_OK;
    // each value as a hexadecimal float literal: exactly the float the block holds
    std::string src = "// the parameter block as constants (pt_fold_params)\nstatic constexpr float PT_UC[" +
                      std::to_string(c->params.size() ? c->params.size() : 1) + "] = {";
    char buf[48];
    for (size_t i = 0; i < c->params.size(); ++i) {
        if (!std::isfinite(c->params[i])) return fail(c, PT_ERR_ARG, "parameter %zu is not finite", i);
        snprintf(buf, sizeof buf, "%s%af", i ? ", " : "", (double)c->params[i]);
        src += buf;
    }
    if (c->params.empty()) src += "0.0f";
    src += "};\n#undef PT_SCENE_PARAMS\n#define PT_SCENE_PARAMS(U) ((U) = (ParamPtr)PT_UC)\n";
    src += c->sceneSrc;
    c->building = true;
    hipModule_t mod = nullptr;
    hipFunction_t fns[K_COUNT];
    int rc = loadSceneModule(c, src.c_str(), mod, fns);
    c->building = false;
    if (rc) return rc;
    hipStreamSynchronize(c->stream);
    destroyGraph(c);
    c->foldedModule = mod;
    memcpy(c->fn, fns, sizeof fns);
    return PT_OK;
}

int pt_set_uniforms(pt_ctx* c, const PtUniforms* u, size_t size) {
    if (!c || !u) return fail(c, PT_ERR_ARG, "uniforms is NULL");
    if (size != sizeof(PtUniforms)) return fail(c, PT_ERR_ARG, "uniforms: %zu bytes, expected %zu", size, sizeof(PtUniforms));
    if (u->width == 0 || u->height == 0 || u->width > 16384 || u->height > 16384) return fail(c, PT_ERR_ARG, "size %u x %u out of range", u->width, u->height);
    if (u->maxbounce > 250) return fail(c, PT_ERR_ARG, "maxbounce %u above 250 (a path counts its bounces in 8 bits)", u->maxbounce);
    if (u->marchRefill > 64) return fail(c, PT_ERR_ARG, "marchRefill %u not in 0..64 (0: the tuning's)", u->marchRefill);
    if (!(u->stepscale > 0.0f && u->stepscale <= 1.0f)) return fail(c, PT_ERR_ARG, "stepscale %g not in (0, 1]", u->stepscale);
    if (!(u->relax == 0.0f || (u->relax >= 1.0f && u->relax <= 1.95f))) return fail(c, PT_ERR_ARG, "relax %g not 0 or in [1, 1.95]", u->relax);
    if (u->measureMode > 2u) return fail(c, PT_ERR_ARG, "measureMode %u not 0, 1 (illuminance) or 2 (luminance)", u->measureMode);
    // the thin lens: a radius of 0 is the pinhole; a lens needs a finite radius and a distance in focus in front of it
    if (!(std::isfinite(u->cam.lensR) && u->cam.lensR >= 0.0f)) return fail(c, PT_ERR_ARG, "camera lensR %g must be finite and >= 0", u->cam.lensR);
    if (u->cam.lensR > 0.0f && !(std::isfinite(u->cam.focus) && u->cam.focus > 0.0f))
        return fail(c, PT_ERR_ARG, "camera focus %g must be > 0 with a lens (lensR %g)", u->cam.focus, u->cam.lensR);
    if (u->lt > 1u) return fail(c, PT_ERR_ARG, "lt %u not 0 or 1", u->lt);
    if ((u->sppm || u->lt) && (u->sppmPhotons == 0 || u->sppmPhotons > (1u << 26)))
        return fail(c, PT_ERR_ARG, "sppmPhotons %u not in 1..2^26 (the photons a sample of the map and of light tracing)", u->sppmPhotons);
    if (u->sppm && !(u->sppmRadius > 0.0f && u->sppmAlpha > 0.0f && u->sppmAlpha < 1.0f))
        return fail(c, PT_ERR_ARG, "sppmRadius %g must be > 0 and sppmAlpha %g in (0, 1)", u->sppmRadius, u->sppmAlpha);
    unsigned int photons = (u->sppm || u->lt) ? u->sppmPhotons : 0u;
    int rc;
    if ((rc = ensureBuffers(c, u->width, u->height, photons))) return rc;
    bool resized = u->width != c->host.u.width || u->height != c->host.u.height;
    c->host.u = *u;
    unsigned int N = u->width * u->height;
    raygenTiles(N, u->width, u->height, c->host.tilesX, c->host.tilesY);
    c->host.invTX = 1.0f / float(c->host.tilesX);
    c->host.invTY = 1.0f / float(c->host.tilesY);
    c->host.nrays = N;
    c->host.nphotons = photons;
    applyRefills(c);
    if (u->guide && (rc = ensureGuide(c))) return rc;
    if (resized) return pt_clear_accum(c);
    return PT_OK;
}

int pt_clear_accum(pt_ctx* c) {
    if (!c) return fail(c, PT_ERR_ARG, "ctx is NULL");
    HIPCHK(c, hipMemsetAsync(c->host.frame, 0, size_t(c->pixels) * 16, c->stream));
    HIPCHK(c, hipMemsetAsync(c->host.accum, 0, size_t(c->pixels) * 32, c->stream));
    HIPCHK(c, hipMemsetAsync(c->host.moments, 0, size_t(c->pixels) * 16, c->stream));
    // every counter but the guide network's step count (its training outlives the picture)
    HIPCHK(c, hipMemsetAsync(c->host.counters, 0, PT_C_GUIDE_T * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(c->host.counters + PT_C_GUIDE_T + 1, 0, (PT_C_WORDS - PT_C_GUIDE_T - 1) * 4, c->stream));
    // a new picture: the meter sets its exposure afresh at the next development
    if (c->aeState) HIPCHK(c, hipMemsetAsync(c->aeState, 0, 8 * 4, c->stream));
    return PT_OK;
}

int pt_trace(pt_ctx* c, uint32_t samples) {
    if (!c) return fail(c, PT_ERR_ARG, "ctx is NULL");
    if (c->building) return fail(c, PT_ERR_STATE, "trace during a scene build");
    if (!c->module) return fail(c, PT_ERR_STATE, "no scene loaded");
    int rc;
    if ((rc = syncDev(c)) || (rc = buildGraph(c))) return rc;
    for (uint32_t s = 0; s < samples; ++s) HIPCHK(c, launchSample(c));
    return PT_OK;
}

int pt_get_tuning(pt_ctx* c, PtTuning* out, size_t size) {
    if (!c || !out) return fail(c, PT_ERR_ARG, "tuning is NULL");
    if (size != sizeof(PtTuning)) return fail(c, PT_ERR_ARG, "tuning: %zu bytes, expected %zu", size, sizeof(PtTuning));
    *out = c->tune;
    return PT_OK;
}

namespace {
int setTuning(pt_ctx* c, unsigned int bpc, unsigned int rm, unsigned int ro) {
    c->tune.blocksPerCu = bpc; c->tune.refillMarch = rm; c->tune.refillOcclude = ro;
    c->blocksPerCu = bpc;
    c->host.gridBlocks = c->cus * bpc;     // part of the graph's key: the next trace captures it anew
    applyRefills(c);
    return PT_OK;
}
// the time a sample of the current setting takes: one sample to build and warm, then `samples` timed on the stream
int timeSetting(pt_ctx* c, uint32_t samples, double& ms) {
    int rc;
    if ((rc = syncDev(c)) || (rc = buildGraph(c))) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(c, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) { hipEventDestroy(e0); return fail(c, PT_ERR_HIP, "event create"); }
    hipError_t e = launchSample(c);
    if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
    for (uint32_t k = 0; e == hipSuccess && k < samples; ++k) e = launchSample(c);
    if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    hipEventDestroy(e0); hipEventDestroy(e1);
    if (e != hipSuccess) return fail(c, PT_ERR_HIP, "autotune: %s", hipGetErrorString(e));
    ms = double(t) / samples;
    return PT_OK;
}
}  // namespace

int pt_set_tuning(pt_ctx* c, const PtTuning* t, size_t size) {
    if (!c || !t) return fail(c, PT_ERR_ARG, "tuning is NULL");
    if (size != sizeof(PtTuning)) return fail(c, PT_ERR_ARG, "tuning: %zu bytes, expected %zu", size, sizeof(PtTuning));
    if (t->blocksPerCu < 1 || t->blocksPerCu > 32 || t->refillMarch < 1 || t->refillMarch > 64 || t->refillOcclude < 1 || t->refillOcclude > 64)
        return fail(c, PT_ERR_ARG, "tuning: blocksPerCu %u not in 1..32 or a refill (%u, %u) not in 1..64", t->blocksPerCu,
                    t->refillMarch, t->refillOcclude);
    return setTuning(c, t->blocksPerCu, t->refillMarch, t->refillOcclude);
}

int pt_autotune(pt_ctx* c, uint32_t samples, PtTuning* out, size_t size) {
    if (!c) return fail(c, PT_ERR_ARG, "ctx is NULL");
    if (out && size != sizeof(PtTuning)) return fail(c, PT_ERR_ARG, "tuning: %zu bytes, expected %zu", size, sizeof(PtTuning));
    if (c->building) return fail(c, PT_ERR_STATE, "autotune during a scene build");
    if (!c->module) return fail(c, PT_ERR_STATE, "no scene loaded");
    if (samples < 1 || samples > 64) return fail(c, PT_ERR_ARG, "samples %u not in 1..64", samples);
    // coordinate descent: the grid of the persistent kernels first (occupancy against the tail of each round), then
    // each kernel's refill threshold (how many idle lanes a wave carries before it takes new rays)
    static const unsigned int kBpc[] = {2, 4, 6, 8, 12, 16};
    static const unsigned int kRefill[] = {1, 4, 8, 16, 24, 32, 48};
    PtTuning best = c->tune;
    unsigned int n = 0;
    int rc;
    double ms = 0.0;
    if ((rc = timeSetting(c, samples, ms))) return rc;
    const double before = ms;
    double bestMs = ms;
    ++n;
    auto tryOne = [&](unsigned int bpc, unsigned int rm, unsigned int ro) -> int {
        if (bpc == best.blocksPerCu && rm == best.refillMarch && ro == best.refillOcclude) return PT_OK;
        int r;
        double t = 0.0;
        if ((r = setTuning(c, bpc, rm, ro)) || (r = timeSetting(c, samples, t))) return r;
        ++n;
        if (t < bestMs) { bestMs = t; best.blocksPerCu = bpc; best.refillMarch = rm; best.refillOcclude = ro; }
        return PT_OK;
    };
    for (unsigned int b : kBpc) if ((rc = tryOne(b, best.refillMarch, best.refillOcclude))) return rc;
    for (unsigned int r : kRefill) if ((rc = tryOne(best.blocksPerCu, r, best.refillOcclude))) return rc;
    for (unsigned int r : kRefill) if ((rc = tryOne(best.blocksPerCu, best.refillMarch, r))) return rc;
    if ((rc = setTuning(c, best.blocksPerCu, best.refillMarch, best.refillOcclude))) return rc;
    c->tune.candidates = n;
    c->tune.msPerSample = bestMs;
    c->tune.msBefore = before;
    if (out) *out = c->tune;
    // the tuning's samples went into the accumulators: the picture starts again
    return pt_clear_accum(c);
}

int pt_sync(pt_ctx* c) {
    if (!c) return fail(c, PT_ERR_ARG, "ctx is NULL");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PT_OK;
}

int pt_read_accum(pt_ctx* c, double* accum, double* moments, size_t npx) {
    if (!c) return fail(c, PT_ERR_ARG, "ctx is NULL");
    size_t px = size_t(c->host.u.width) * c->host.u.height;
    if (npx != px) return fail(c, PT_ERR_ARG, "npx %zu, the image has %zu pixels", npx, px);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (accum) HIPCHK(c, hipMemcpy(accum, c->host.accum, px * 32, hipMemcpyDeviceToHost));
    if (moments) HIPCHK(c, hipMemcpy(moments, c->host.moments, px * 16, hipMemcpyDeviceToHost));
    return PT_OK;
}

int pt_read_stats(pt_ctx* c, uint64_t* stats, size_t n) {
    if (!c || !stats || n > PT_S_WORDS) return fail(c, PT_ERR_ARG, "stats: NULL or n > %d", PT_S_WORDS);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(stats, c->host.stats, n * 8, hipMemcpyDeviceToHost));
    return PT_OK;
}

int pt_reset_stats(pt_ctx* c) {
    if (!c) return fail(c, PT_ERR_ARG, "ctx is NULL");
    HIPCHK(c, hipMemsetAsync(c->host.stats, 0, PT_S_WORDS * 8, c->stream));
    return PT_OK;
}

int pt_time_kernels(pt_ctx* c, uint32_t samples, double* ms, size_t n) {
    if (!c || !ms || n < 5) return fail(c, PT_ERR_ARG, "ms: NULL or fewer than 5 slots");
    if (!c->module) return fail(c, PT_ERR_STATE, "no scene loaded");
    int rc;
    if ((rc = syncDev(c))) return rc;
    for (size_t k = 0; k < n; ++k) ms[k] = 0.0;
    // a sample's launches at most: the photon pass 2 + 10 a round (with light tracing's binning of its connections),
    // the camera's 1 + 11 a round (guide inference included), flush and resolve, collect and 3 a training step
    size_t launches = 6 + 21 * size_t(rounds(c)) + 3 * size_t(kGuideMaxSteps);
    std::vector<hipEvent_t> ev(2 * launches);
    for (auto& e : ev) HIPCHK(c, hipEventCreate(&e));
    for (uint32_t s = 0; s < samples; ++s) {
        std::vector<int> kinds;
        hipError_t e = enqueueSample(c, &ev, &kinds);
        if (e != hipSuccess) { for (auto& x : ev) hipEventDestroy(x); return fail(c, PT_ERR_HIP, "launch: %s", hipGetErrorString(e)); }
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < kinds.size(); ++i) {
            float t = 0.0f;
            hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]);
            int k = kinds[i];
            int slot = k == K_GENRAYS ? 0 : k == K_MARCH ? 1 : k == K_SHADE ? 2 : k == K_OCCLUDE ? 3 : (k == K_RESOLVE || k == K_FLUSH) ? 4
                     : (k == K_SPPMCLEAR || k == K_PHOTONGEN) ? 6 : k == K_SURFACE ? 7
                     : (k == K_GINFER || k == K_GCOLLECT || k == K_GGRAD || k == K_GADAM || k == K_GCOMMIT) ? 8 : 5;
            if (slot < (int)n) ms[slot] += t;
            // the guide's kernels one by one: inference, the gradient pass (forward and backward), the rest of training
            int g = k == K_GINFER ? 9 : k == K_GGRAD ? 10 : (k == K_GCOLLECT || k == K_GADAM || k == K_GCOMMIT) ? 11 : -1;
            if (g >= 0 && g < (int)n) ms[g] += t;
        }
    }
    for (auto& e : ev) hipEventDestroy(e);
    return PT_OK;
}

// ---- the optics trace over the loaded scene (optics/index.js createOptics run / readHits / results)
namespace {
bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }
double len3(const float* v) { return std::sqrt(double(v[0]) * v[0] + double(v[1]) * v[1] + double(v[2]) * v[2]); }
// a unit vector square to `a` (unit), from `u` where it is not along `a`
void squareTo(const float* a, const float* u, float* out) {
    double v[3] = {u[0], u[1], u[2]};
    double d = v[0] * a[0] + v[1] * a[1] + v[2] * a[2];
    for (int k = 0; k < 3; ++k) v[k] -= d * a[k];
    double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (!(l > 1e-3)) {
        // any tangent: the axis crossed with the world axis it is least along
        double w[3] = {0, 0, 0};
        w[std::fabs(a[0]) < 0.9 ? 0 : 1] = 1.0;
        v[0] = a[1] * w[2] - a[2] * w[1]; v[1] = a[2] * w[0] - a[0] * w[2]; v[2] = a[0] * w[1] - a[1] * w[0];
        l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    for (int k = 0; k < 3; ++k) out[k] = float(v[k] / l);
}
hipError_t launchOpt(pt_ctx* c, int k, unsigned int grid, bool withQ, unsigned int qsel) {
    void* args2[] = {&c->dev, &c->optDev};
    void* args3[] = {&c->dev, &c->optDev, &qsel};
    return hipModuleLaunchKernel(c->fn[k], grid, 1, 1, 256, 1, 1, 0, c->stream, withQ ? args3 : args2, nullptr);
}
}  // namespace

int pt_optics_run(pt_ctx* c, const PtOpticsSource* src, uint32_t nsrc, const PtOpticsTrace* t, size_t tsize,
                  PtOpticsTally* out, size_t osize, double* far, size_t nfar, float* hits, uint32_t hitCap) {
    if (!c) return fail(c, PT_ERR_ARG, "ctx is NULL");
    if (!t || !out || (nsrc && !src)) return fail(c, PT_ERR_ARG, "sources, trace or tally is NULL");
    if (tsize != sizeof(PtOpticsTrace)) return fail(c, PT_ERR_ARG, "trace: %zu bytes, expected %zu", tsize, sizeof(PtOpticsTrace));
    if (osize != sizeof(PtOpticsTally)) return fail(c, PT_ERR_ARG, "tally: %zu bytes, expected %zu", osize, sizeof(PtOpticsTally));
    if (c->building) return fail(c, PT_ERR_STATE, "optics trace during a scene build");
    if (!c->module) return fail(c, PT_ERR_STATE, "no scene loaded");
    if (t->rays == 0) return fail(c, PT_ERR_ARG, "rays: 0");
    if (t->bounces > 250) return fail(c, PT_ERR_ARG, "bounces %u above 250 (a ray counts its bounces in 8 bits)", t->bounces);
    if (!(std::isfinite(t->maxDist) && t->maxDist >= 0.0f)) return fail(c, PT_ERR_ARG, "maxDist %g must be finite and >= 0", t->maxDist);
    if (t->farMode > 2u) return fail(c, PT_ERR_ARG, "farMode %u not 0 (off), 1 (escaping) or 2 (a detector)", t->farMode);
    if (t->marchSteps > 60000u) return fail(c, PT_ERR_ARG, "marchSteps %u above 60000 (a ray counts its steps in 16 bits)", t->marchSteps);
    const size_t bins = t->farMode ? size_t(t->nC) * t->nG : 0;
    if (t->farMode) {
        if (t->nC < 1 || t->nC > 3600 || t->nG < 2 || t->nG > 1801) return fail(c, PT_ERR_ARG, "far field %u x %u: nC 1..3600, nG 2..1801", t->nC, t->nG);
        if (!far || nfar != bins) return fail(c, PT_ERR_ARG, "far: %zu bins given, the grid has %zu", far ? nfar : size_t(0), bins);
        if (t->farMode == 2u && t->farMaterial >= PT_MAX_MATERIALS) return fail(c, PT_ERR_ARG, "farMaterial %u not below %d", t->farMaterial, PT_MAX_MATERIALS);
        if (!finite3(t->farAxis) || !(len3(t->farAxis) > 0.0) || !finite3(t->farC0)) return fail(c, PT_ERR_ARG, "far axis or C0 not a direction");
    }
    const unsigned int detCap = hits ? std::min(t->detCap, hitCap) : 0u;
    // the sources with flux, checked; their hues at luminance 1; their cumulative shares
    std::vector<PtOpticsSource> live;
    double total = 0.0;
    for (uint32_t i = 0; i < nsrc; ++i) {
        PtOpticsSource q = src[i];
        if (!(std::isfinite(q.flux) && q.flux > 0.0f)) continue;
        if (q.shape > 2u || q.distribution > 2u) return fail(c, PT_ERR_ARG, "source %u: shape %u (0..2) or distribution %u (0..2)", i, q.shape, q.distribution);
        if (!finite3(q.pos) || !finite3(q.axis) || !(len3(q.axis) > 0.0) || !finite3(q.u))
            return fail(c, PT_ERR_ARG, "source %u: position, axis or tangent not finite (or no axis)", i);
        if (!(std::isfinite(q.halfAngle) && std::isfinite(q.n) && q.n >= 0.0f && std::isfinite(q.size[0]) && q.size[0] >= 0.0f &&
              std::isfinite(q.size[1]) && q.size[1] >= 0.0f))
            return fail(c, PT_ERR_ARG, "source %u: half angle, n or size not finite (or negative)", i);
        const double al = len3(q.axis);
        for (int k = 0; k < 3; ++k) q.axis[k] = float(q.axis[k] / al);
        squareTo(q.axis, q.u, q.u);
        double Y = 0.2126 * q.colour[0] + 0.7152 * q.colour[1] + 0.0722 * q.colour[2];
        if (!finite3(q.colour) || !(Y > 0.0)) { q.colour[0] = q.colour[1] = q.colour[2] = 1.0f; Y = 1.0; }
        for (int k = 0; k < 3; ++k) q.colour[k] = float(q.colour[k] / Y);
        if (!(q.emitN >= 1.0f) || !std::isfinite(q.emitN)) q.emitN = 1.0f;
        live.push_back(q);
        total += q.flux;
    }
    if (live.empty()) return fail(c, PT_ERR_ARG, "no sources: none with flux");
    std::vector<float> cdf(live.size());
    double cum = 0.0;
    for (size_t i = 0; i < live.size(); ++i) { cum += live[i].flux; cdf[i] = i + 1 == live.size() ? 1.0f : float(cum / total); }
    // the device table and its buffers
    int rc;
    if (!c->optDev && (rc = allocBuf(c, c->optDev, sizeof(PtOptDev)))) return rc;
    if (!c->opt.tally && (rc = allocBuf(c, c->opt.tally, PT_OT_WORDS * 8))) return rc;
    if (live.size() > c->optSrcCap) {
        void* p = (void*)c->opt.src; freeBuf(c, p);
        p = (void*)c->opt.cdf; freeBuf(c, p);
        PtOpticsSource* ps = nullptr; float* pc = nullptr;
        if ((rc = allocBuf(c, ps, live.size() * sizeof(PtOpticsSource))) || (rc = allocBuf(c, pc, live.size() * 4))) return rc;
        c->opt.src = ps; c->opt.cdf = pc;
        c->optSrcCap = (unsigned int)live.size();
    }
    if (bins > c->optFarCap) {
        void* p = c->opt.far; freeBuf(c, p); c->opt.far = nullptr;
        if ((rc = allocBuf(c, c->opt.far, bins * 8))) return rc;
        c->optFarCap = (unsigned int)bins;
    }
    if (detCap > c->optHitCap) {
        void* p = c->opt.hits; freeBuf(c, p); c->opt.hits = nullptr;
        if ((rc = allocBuf(c, c->opt.hits, size_t(detCap) * PT_OPT_HIT_WORDS * 4))) return rc;
        c->optHitCap = detCap;
    }
    // a batch as many rays as there are slots (at least 2^20); the draws are keyed by the ray's id, so batching does
    // not change which rays are traced
    const unsigned int batch = std::min(t->rays, std::max(c->capacity, 1u << 20));
    // the trace's own step budget, through the engine's maxiter while it runs (restored after, whatever happens)
    struct KeepMaxiter {
        pt_ctx* c; unsigned int v;
        ~KeepMaxiter() { if (c->host.u.maxiter != v) { c->host.u.maxiter = v; c->devDirty = true; } }
    } keep{c, c->host.u.maxiter};
    const unsigned int steps = t->marchSteps ? t->marchSteps : 2000u;
    if (c->host.u.maxiter != steps) { c->host.u.maxiter = steps; c->devDirty = true; }
    if ((rc = ensureSlots(c, batch)) || (rc = syncDev(c))) return rc;
    PtOptDev& O = c->opt;
    O.nsrc = (unsigned int)live.size();
    O.bounces = t->bounces;
    O.t = *t;
    O.t.detCap = detCap;
    if (t->farMode) {
        const double al = len3(t->farAxis);
        for (int k = 0; k < 3; ++k) O.t.farAxis[k] = float(t->farAxis[k] / al);
        squareTo(O.t.farAxis, t->farC0, O.t.farC0);
    }
    // the trace's seed through a 64-bit mix (splitmix64), four words
    uint64_t z = 0x9E3779B97F4A7C15ull * (uint64_t(t->seed) + 1u);
    for (int k = 0; k < 4; ++k) {
        z += 0x9E3779B97F4A7C15ull;
        uint64_t x = z;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        O.seed[k] = uint32_t(x ^ (x >> 31));
    }
    HIPCHK(c, hipMemcpyAsync((void*)O.src, live.data(), live.size() * sizeof(PtOpticsSource), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync((void*)O.cdf, cdf.data(), cdf.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(O.tally, 0, PT_OT_WORDS * 8, c->stream));
    if (bins) HIPCHK(c, hipMemsetAsync(O.far, 0, bins * 8, c->stream));
    const unsigned int R = t->bounces + 2u;     // the first segment, one per interaction, one for hits from behind
    for (unsigned int first = 0; first < t->rays; first += batch) {
        O.rays = std::min(batch, t->rays - first);
        O.firstId = first;
        HIPCHK(c, hipMemcpyAsync(c->optDev, &O, sizeof(PtOptDev), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, launchOpt(c, K_OPTGEN, (O.rays + 255u) / 256u, false, 0));
        for (unsigned int it = 0; it < R; ++it) {
            const unsigned int in = it & 1u;
            HIPCHK(c, launch(c, K_BINCOUNT, gridFor(c, K_BINCOUNT), in));
            HIPCHK(c, launch(c, K_BINSCAN, gridFor(c, K_BINSCAN), 0));
            HIPCHK(c, launch(c, K_BINSCATTER, gridFor(c, K_BINSCATTER), in));
            HIPCHK(c, launch(c, K_MARCH, gridFor(c, K_MARCH), in));
            HIPCHK(c, launch(c, K_SURFACE, gridFor(c, K_SURFACE), in ^ 1u));
            HIPCHK(c, launchOpt(c, K_OPTSHADE, c->host.gridBlocks, true, in ^ 1u));
            HIPCHK(c, launch(c, K_OCCLUDE, gridFor(c, K_OCCLUDE), in));
        }
        HIPCHK(c, launchOpt(c, K_OPTFLUSH, c->host.gridBlocks, true, R & 1u));
        // the table is rewritten for the next batch: this one's kernels finish first
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    std::vector<unsigned long long> T(PT_OT_WORDS);
    HIPCHK(c, hipMemcpy(T.data(), O.tally, PT_OT_WORDS * 8, hipMemcpyDeviceToHost));
    // weights to lumens: every ray carries the same share of the total flux
    const double k = total / double(t->rays) / PT_OPT_FX;
    memset(out, 0, sizeof *out);
    out->emitted = total;
    out->escaped = T[PT_OT_ESCAPED] * k;
    out->lost = T[PT_OT_LOST] * k;
    out->detected = T[PT_OT_DETECTED] * k;
    out->volume = T[PT_OT_VOLUME] * k;
    out->rays = t->rays;
    out->escapedRays = T[PT_OT_CESC];
    out->lostRays = T[PT_OT_CLOST];
    out->interactions = T[PT_OT_CINTER];
    out->absorbedRays = T[PT_OT_CABS];
    out->detectedRays = T[PT_OT_CDET];
    out->detectorHits = T[PT_OT_HITS];
    double taken = 0.0;
    for (int m = 0; m < PT_MAX_MATERIALS; ++m) {
        out->incident[m] = T[PT_OT_MAT + 3 * m] * k;
        out->leaving[m] = T[PT_OT_MAT + 3 * m + 1] * k;
        out->hits[m] = T[PT_OT_MAT + 3 * m + 2];
        taken += out->incident[m] - out->leaving[m];
    }
    out->absorbed = taken - out->detected + out->volume;
    if (bins) {
        std::vector<unsigned long long> F(bins);
        HIPCHK(c, hipMemcpy(F.data(), O.far, bins * 8, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < bins; ++i) far[i] = F[i] * k;
    }
    const unsigned int nh = (unsigned int)std::min<unsigned long long>(T[PT_OT_HITS], detCap);
    if (nh) {
        HIPCHK(c, hipMemcpy(hits, O.hits, size_t(nh) * PT_OPT_HIT_WORDS * 4, hipMemcpyDeviceToHost));
        const float kf = float(total / double(t->rays));
        for (unsigned int i = 0; i < nh; ++i) hits[i * PT_OPT_HIT_WORDS + 3] *= kf;
    }
    return PT_OK;
}

// The settings checked, the buffers made, the settings on the device and, with autoExposure, the picture metered.
static int developBegin(pt_ctx* c, const PtDevelop* d, size_t dsize) {
    if (!d) return fail(c, PT_ERR_ARG, "settings is NULL");
    if (dsize != sizeof(PtDevelop)) return fail(c, PT_ERR_ARG, "develop: %zu bytes, expected %zu", dsize, sizeof(PtDevelop));
    if (c->building) return fail(c, PT_ERR_STATE, "develop during a scene build");
    if (!c->module) return fail(c, PT_ERR_STATE, "no scene loaded");
    if (d->toneMap > 7u) return fail(c, PT_ERR_ARG, "toneMap %u not in 0..7", d->toneMap);
    if (d->transfer > 3u) return fail(c, PT_ERR_ARG, "transfer %u not in 0..3", d->transfer);
    if (d->measureMode > 2u) return fail(c, PT_ERR_ARG, "measureMode %u not in 0..2", d->measureMode);
    if (d->lightOn & ~PT_LIGHT_ALL) return fail(c, PT_ERR_ARG, "lightOn 0x%x has bits beyond 0x%x", d->lightOn, PT_LIGHT_ALL);
    const float scalars[] = {d->exposureEV, d->aeSpeed, d->aeHeadroom, d->contrast, d->saturation, d->brightness, d->measureMin,
                             d->measureMax};
    for (float v : scalars)
        if (!std::isfinite(v)) return fail(c, PT_ERR_ARG, "exposureEV, aeSpeed, aeHeadroom, the grading and the measure scale must be finite");
    for (const float* v = d->wb; v < d->pad; ++v)
        if (!std::isfinite(*v)) return fail(c, PT_ERR_ARG, "a darkroom setting (word %d) is not finite", int(v - d->wb));
    if (d->lightOn & PT_LIGHT_CURVES)
        for (int k = 0; k < 4 * PT_CURVE_LUT; ++k)
            if (!std::isfinite(d->curveLut[k])) return fail(c, PT_ERR_ARG, "curveLut[%d] is not finite", k);
    if (d->measureMode && !(d->measureMin > 0.0f && d->measureMax > 0.0f))
        return fail(c, PT_ERR_ARG, "the false colour's scale needs measureMin, measureMax > 0");
    if ((d->lightOn & PT_LIGHT_GRAIN) && !(d->grain[1] >= 1.0f)) return fail(c, PT_ERR_ARG, "grain size %g below 1 pixel", d->grain[1]);
    if ((d->lightOn & PT_LIGHT_LEVELS) && !(d->levels[2] > 0.0f)) return fail(c, PT_ERR_ARG, "levels 1/gamma %g not > 0", d->levels[2]);
    if (d->toneMap == 7u && !(d->tone7[0] > 0.0f && d->tone7[1] >= 0.0f && d->tone7[2] >= 0.0f))
        return fail(c, PT_ERR_ARG, "curve 7 needs white > 0, toe and shoulder >= 0");
    int rc;
    if (!c->aeHist) {
        if ((rc = allocBuf(c, c->aeHist, 128 * 4)) || (rc = allocBuf(c, c->aeState, 8 * 4)) || (rc = allocBuf(c, c->devSet, sizeof(PtDevelop))))
            return rc;
        HIPCHK(c, hipMemsetAsync(c->aeHist, 0, 128 * 4, c->stream));
        HIPCHK(c, hipMemsetAsync(c->aeState, 0, 8 * 4, c->stream));
    }
    if ((rc = syncDev(c))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->devSet, d, sizeof(PtDevelop), hipMemcpyHostToDevice, c->stream));
    if (d->autoExposure) {
        void* ah[] = {&c->dev, &c->aeHist};
        HIPCHK(c, hipModuleLaunchKernel(c->fn[K_AEHIST], c->cus * 4u, 1, 1, 256, 1, 1, 0, c->stream, ah, nullptr));
        void* au[] = {&c->aeHist, &c->aeState, &c->devSet};
        HIPCHK(c, hipModuleLaunchKernel(c->fn[K_AEUPDATE], 1, 1, 1, 64, 1, 1, 0, c->stream, au, nullptr));
    }
    return PT_OK;
}

static int developInfo(pt_ctx* c, const PtDevelop* d, PtDevelopInfo* info) {
    if (!info) return PT_OK;
    float st[8];
    HIPCHK(c, hipMemcpy(st, c->aeState, sizeof st, hipMemcpyDeviceToHost));
    info->ev = d->autoExposure ? st[0] : d->exposureEV;
    info->log2Average = st[2];
    info->log2High = st[3];
    info->metered = d->autoExposure ? st[4] : 0.0f;
    return PT_OK;
}

int pt_develop(pt_ctx* c, const PtDevelop* d, size_t dsize, uint8_t* rgba, size_t npx, PtDevelopInfo* info) {
    if (!c) return fail(c, PT_ERR_ARG, "ctx is NULL");
    if (!rgba) return fail(c, PT_ERR_ARG, "rgba is NULL");
    const size_t px = size_t(c->host.u.width) * c->host.u.height;
    if (c->module && npx != px) return fail(c, PT_ERR_ARG, "npx %zu, the picture has %zu pixels", npx, px);
    int rc;
    if ((rc = developBegin(c, d, dsize))) return rc;
    if (px > c->devRgbaPx) {
        void* p = c->devRgba; freeBuf(c, p); c->devRgba = nullptr;
        if ((rc = allocBuf(c, c->devRgba, px * 4))) return rc;
        c->devRgbaPx = px;
    }
    void* dv[] = {&c->dev, &c->devRgba, &c->aeState, &c->devSet};
    HIPCHK(c, hipModuleLaunchKernel(c->fn[K_DEVELOP], (unsigned int)((px + 255) / 256), 1, 1, 256, 1, 1, 0, c->stream, dv, nullptr));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(rgba, c->devRgba, px * 4, hipMemcpyDeviceToHost));
    return developInfo(c, d, info);
}

int pt_develop16(pt_ctx* c, const PtDevelop* d, size_t dsize, uint16_t* rgba16, size_t npx, PtDevelopInfo* info) {
    if (!c) return fail(c, PT_ERR_ARG, "ctx is NULL");
    if (!rgba16) return fail(c, PT_ERR_ARG, "rgba16 is NULL");
    if (!d || dsize != sizeof(PtDevelop)) return developBegin(c, d, dsize);
    const size_t px = size_t(c->host.u.width) * c->host.u.height;
    if (c->module && npx != px) return fail(c, PT_ERR_ARG, "npx %zu, the picture has %zu pixels", npx, px);
    // 16 bits keep what the stages computed: no dither (it breaks 8-bit steps, and there are none); the rest as asked
    PtDevelop s = *d;
    s.lightOn &= ~PT_LIGHT_DITHER;
    int rc;
    if ((rc = developBegin(c, &s, dsize))) return rc;
    uint16_t* dout = nullptr;
    if ((rc = allocBuf(c, dout, px * 8))) return rc;
    void* dv[] = {&c->dev, &dout, &c->aeState, &c->devSet};
    hipError_t e = hipModuleLaunchKernel(c->fn[K_DEVELOP16], (unsigned int)((px + 255) / 256), 1, 1, 256, 1, 1, 0, c->stream, dv, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(rgba16, dout, px * 8, hipMemcpyDeviceToHost);
    void* p = dout;
    freeBuf(c, p);
    HIPCHK(c, e);
    return developInfo(c, &s, info);
}

int pt_develop_view(pt_ctx* c, const PtDevelop* d, size_t dsize, const float* view, uint32_t outW, uint32_t outH, uint32_t clipWarn,
                    uint8_t* rgba, size_t npx, PtDevelopInfo* info) {
    if (!c) return fail(c, PT_ERR_ARG, "ctx is NULL");
    if (!rgba || !view) return fail(c, PT_ERR_ARG, "view or rgba is NULL");
    if (outW == 0 || outH == 0 || outW > 16384 || outH > 16384) return fail(c, PT_ERR_ARG, "output %u x %u not in 1..16384", outW, outH);
    if (npx != size_t(outW) * outH) return fail(c, PT_ERR_ARG, "npx %zu, the output has %zu pixels", npx, size_t(outW) * outH);
    for (int k = 0; k < 4; ++k)
        if (!std::isfinite(view[k])) return fail(c, PT_ERR_ARG, "view[%d] is not finite", k);
    if (!(view[2] > 0.0f && view[3] > 0.0f)) return fail(c, PT_ERR_ARG, "the view's span (%g, %g) must be > 0", view[2], view[3]);
    int rc;
    if ((rc = developBegin(c, d, dsize))) return rc;
    // taps a side: the larger span rounded up, 1 at full size or closer, at most 8
    const uint32_t box = (uint32_t)std::min(8.0, std::max(1.0, std::ceil(std::max(double(view[2]), double(view[3])) - 1e-9)));
    const size_t px = size_t(outW) * outH;
    uint8_t* dout = nullptr;
    if ((rc = allocBuf(c, dout, px * 4))) return rc;
    float4 v = make_float4(view[0], view[1], view[2], view[3]);
    void* vv[] = {&c->dev, &dout, &c->aeState, &c->devSet, &v, &outW, &outH, (void*)&box, &clipWarn};
    hipError_t e = hipModuleLaunchKernel(c->fn[K_DEVVIEW], (unsigned int)((px + 255) / 256), 1, 1, 256, 1, 1, 0, c->stream, vv, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(rgba, dout, px * 4, hipMemcpyDeviceToHost);
    void* p = dout;
    freeBuf(c, p);
    HIPCHK(c, e);
    return developInfo(c, d, info);
}

int pt_develop_histogram(pt_ctx* c, const PtDevelop* d, size_t dsize, uint32_t step, uint32_t* hist, size_t nbins, PtDevelopInfo* info) {
    if (!c) return fail(c, PT_ERR_ARG, "ctx is NULL");
    if (!hist) return fail(c, PT_ERR_ARG, "hist is NULL");
    if (nbins != 4 * PT_HIST_BINS) return fail(c, PT_ERR_ARG, "nbins %zu, the histogram has %d", nbins, 4 * PT_HIST_BINS);
    if (step == 0) return fail(c, PT_ERR_ARG, "step 0");
    if (d && d->measureMode) return fail(c, PT_ERR_ARG, "a false-colour view is a measurement: no histogram of it");
    int rc;
    if ((rc = developBegin(c, d, dsize))) return rc;
    uint32_t* dh = nullptr;
    if ((rc = allocBuf(c, dh, 4 * PT_HIST_BINS * 4))) return rc;
    const size_t W = c->host.u.width, H = c->host.u.height;
    const size_t n = ((W + step - 1) / step) * ((H + step - 1) / step);
    const unsigned int blocks = (unsigned int)std::min<size_t>((n + 255) / 256, size_t(c->cus) * 4u);
    void* hv[] = {&c->dev, &dh, &c->aeState, &c->devSet, &step};
    hipError_t e = hipMemsetAsync(dh, 0, 4 * PT_HIST_BINS * 4, c->stream);
    if (e == hipSuccess)
        e = hipModuleLaunchKernel(c->fn[K_DEVHIST], std::max(blocks, 1u), 1, 1, 256, 1, 1, 0, c->stream, hv, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(hist, dh, 4 * PT_HIST_BINS * 4, hipMemcpyDeviceToHost);
    void* p = dh;
    freeBuf(c, p);
    HIPCHK(c, e);
    return developInfo(c, d, info);
}

int pt_pick_rays(pt_ctx* c, float* pos4, uint32_t* info, size_t n) {
    if (!c || !pos4 || !info) return fail(c, PT_ERR_ARG, "pos4 or info is NULL");
    if (!c->module) return fail(c, PT_ERR_STATE, "no scene loaded");
    if (n != c->host.nrays) return fail(c, PT_ERR_ARG, "n %zu, a sample has %u camera rays", n, c->host.nrays);
    int rc;
    if ((rc = syncDev(c))) return rc;
    float* dpos = nullptr;
    unsigned int* dinfo = nullptr;
    if ((rc = allocBuf(c, dpos, n * 16)) || (rc = allocBuf(c, dinfo, n * 8))) return rc;
    HIPCHK(c, launch(c, K_GENRAYS, gridFor(c, K_GENRAYS), 0));
    HIPCHK(c, launch(c, K_BINCOUNT, gridFor(c, K_BINCOUNT), 0));
    HIPCHK(c, launch(c, K_BINSCAN, gridFor(c, K_BINSCAN), 0));
    HIPCHK(c, launch(c, K_BINSCATTER, gridFor(c, K_BINSCATTER), 0));
    HIPCHK(c, launch(c, K_MARCH, gridFor(c, K_MARCH), 0));
    void* args[] = {&c->dev, &dpos, &dinfo};
    HIPCHK(c, hipModuleLaunchKernel(c->fn[K_PICK], (unsigned int)((n + 255) / 256), 1, 1, 256, 1, 1, 0, c->stream, args, nullptr));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(pos4, dpos, n * 16, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(info, dinfo, n * 8, hipMemcpyDeviceToHost));
    void* p1 = dpos; freeBuf(c, p1);
    void* p2 = dinfo; freeBuf(c, p2);
    return PT_OK;
}

int pt_guide_config(pt_ctx* c, const PtGuideConfig* g, size_t size) {
    if (!c || !g) return fail(c, PT_ERR_ARG, "config is NULL");
    if (size != sizeof(PtGuideConfig)) return fail(c, PT_ERR_ARG, "config: %zu bytes, expected %zu", size, sizeof(PtGuideConfig));
    if (g->batch < 64 || g->batch > (1u << 20) || g->batch % 64) return fail(c, PT_ERR_ARG, "batch %u: a multiple of 64 in [64, 2^20]", g->batch);
    if (g->steps < 1 || g->steps > kGuideMaxSteps) return fail(c, PT_ERR_ARG, "steps %u not in 1..%u", g->steps, kGuideMaxSteps);
    if (!(g->lr > 0.0f)) return fail(c, PT_ERR_ARG, "lr %g must be positive", g->lr);
    for (int a = 0; a < 3; ++a)
        if (!(g->boxMax[a] > g->boxMin[a])) return fail(c, PT_ERR_ARG, "the box is empty on axis %d", a);
    hipStreamSynchronize(c->stream);
    destroyGraph(c);
    c->host.g = *g;
    c->devDirty = true;
    int rc;
    if ((rc = ensureGuide(c))) return rc;
    std::vector<float> th;
    guideInit(g->seed, th);
    return guideUpload(c, th.data());
}

int pt_guide_weights(pt_ctx* c, float* theta, size_t n) {
    if (!c || !theta || n != kGuideNP) return fail(c, PT_ERR_ARG, "theta: NULL or n != %u", kGuideNP);
    if (!c->guideReady) return fail(c, PT_ERR_STATE, "the guide network is not set up (pt_guide_config or uniforms.guide)");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(theta, c->host.gMaster, kGuideNP * 4, hipMemcpyDeviceToHost));
    return PT_OK;
}

int pt_guide_load(pt_ctx* c, const float* theta, size_t n) {
    if (!c || !theta || n != kGuideNP) return fail(c, PT_ERR_ARG, "theta: NULL or n != %u", kGuideNP);
    int rc;
    if (!c->guideReady && (rc = ensureGuide(c))) return rc;
    return guideUpload(c, theta);
}

int pt_guide_forward(pt_ctx* c, const float* raw, size_t n, float* out41) {
    if (!c || !raw || !out41 || n == 0 || n > (1u << 24)) return fail(c, PT_ERR_ARG, "raw/out41 NULL or n not in 1..2^24");
    if (!c->module) return fail(c, PT_ERR_STATE, "no scene loaded");
    int rc;
    if (!c->guideReady && (rc = ensureGuide(c))) return rc;
    if ((rc = syncDev(c))) return rc;
    float *din = nullptr, *dout = nullptr;
    if ((rc = allocBuf(c, din, n * 40)) || (rc = allocBuf(c, dout, n * 164))) return rc;
    HIPCHK(c, hipMemcpy(din, raw, n * 40, hipMemcpyHostToDevice));
    unsigned int nn = (unsigned int)n;
    void* args[] = {&c->dev, &din, &nn, &dout};
    unsigned int grid = (unsigned int)std::min<size_t>((n + 63) / 64, 4096);
    HIPCHK(c, hipModuleLaunchKernel(c->fn[K_GEVAL], grid, 1, 1, 64, 1, 1, 0, c->stream, args, nullptr));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(out41, dout, n * 164, hipMemcpyDeviceToHost));
    void* p1 = din; freeBuf(c, p1);
    void* p2 = dout; freeBuf(c, p2);
    return PT_OK;
}

int pt_guide_train(pt_ctx* c, const float* samples, size_t n, float* grad, double* loss) {
    if (!c || !samples) return fail(c, PT_ERR_ARG, "samples is NULL");
    if (!c->module) return fail(c, PT_ERR_STATE, "no scene loaded");
    int rc;
    if (!c->guideReady && (rc = ensureGuide(c))) return rc;
    if (n == 0 || n > c->host.g.batch || n > c->host.gBatchCap) return fail(c, PT_ERR_ARG, "n %zu not in 1..%u", n, c->host.g.batch);
    if ((rc = syncDev(c))) return rc;
    double sumF = 0.0;
    for (size_t i = 0; i < n; ++i) sumF += samples[16 * i + 7];
    if (!(sumF > 0.0)) return fail(c, PT_ERR_ARG, "the samples' weights sum to %g", sumF);
    const unsigned int cnt = (unsigned int)n;
    const double zero = 0.0;
    float* dgrad = nullptr;
    if (grad && (rc = allocBuf(c, dgrad, kGuideNP * 4))) return rc;
    HIPCHK(c, hipMemcpy(c->host.gBatch, samples, n * 64, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->host.counters + PT_C_GUIDE_N, &cnt, 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->host.gSumF, &sumF, 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->host.gLoss, &zero, 8, hipMemcpyHostToDevice));
    // the gradient's copy goes through the device table for this one step
    c->host.gGradOut = dgrad;
    HIPCHK(c, hipMemcpy(c->dev, &c->host, sizeof(PtDev), hipMemcpyHostToDevice));
    hipError_t e;
    if ((e = launch(c, K_GGRAD, gridFor(c, K_GGRAD), 0)) || (e = launch(c, K_GADAM, gridFor(c, K_GADAM), 0)) ||
        (e = launch(c, K_GCOMMIT, gridFor(c, K_GCOMMIT), 0))) {
        c->host.gGradOut = nullptr;
        c->devDirty = true;
        return fail(c, PT_ERR_HIP, "guide step: %s", hipGetErrorString(e));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->host.gGradOut = nullptr;
    HIPCHK(c, hipMemcpy(c->dev, &c->host, sizeof(PtDev), hipMemcpyHostToDevice));
    if (grad) HIPCHK(c, hipMemcpy(grad, dgrad, kGuideNP * 4, hipMemcpyDeviceToHost));
    if (loss) HIPCHK(c, hipMemcpy(loss, c->host.gLoss, 8, hipMemcpyDeviceToHost));
    if (dgrad) { void* p = dgrad; freeBuf(c, p); }
    return PT_OK;
}

int pt_guide_status(pt_ctx* c, double* out, size_t n) {
    if (!c || !out || n < 2) return fail(c, PT_ERR_ARG, "out: NULL or n < 2");
    if (!c->guideReady) return fail(c, PT_ERR_STATE, "the guide network is not set up");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    unsigned int w[2];
    HIPCHK(c, hipMemcpy(w, c->host.counters + PT_C_GUIDE_N, 8, hipMemcpyDeviceToHost));
    out[0] = w[1];
    out[1] = w[0];
    const size_t k = std::min<size_t>(n - 2, c->host.g.steps);
    if (k) HIPCHK(c, hipMemcpy(out + 2, c->host.gLoss, k * 8, hipMemcpyDeviceToHost));
    return PT_OK;
}

}  // extern "C"

