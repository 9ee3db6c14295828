/* libpt: the C ABI of the MI355X wavefront SDF path tracer, a drop-in for the engine behind ShadyTracer's
 * lib/pathtracer. The JavaScript shim (shadytracer_amd/js/raytrace_main.js) keeps the reference's
 * RAYTRACER(WG, _U, Q, scene) object and calls these through the N-API addon (shadytracer_amd/js/addon.cc);
 * Python reaches them through ctypes (shadytracer_amd/pt.py).
 *
 * Every call returns 0 on success or a PT_ERR_* code; pt_last_error(ctx) then says what went wrong. The
 * library owns all device memory and its stream; readbacks copy into caller-provided arrays.
 * Threading: a context is used from one thread at a time. pt_trace only enqueues; pt_sync waits.
 */
#ifndef PT_H
#define PT_H

#include <stddef.h>
#include <stdint.h>
#include "pt_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PT_OK 0
#define PT_ERR_ARG 1        /* a bad argument: null pointer, size out of range, wrong struct size */
#define PT_ERR_HIP 2        /* the HIP runtime failed (device lost, out of memory, launch failure) */
#define PT_ERR_COMPILE 3    /* the scene did not compile; the previous scene is still loaded */
#define PT_ERR_STATE 4      /* the call is not valid now: trace with no scene, or during a scene build */

#define PT_ABI_VERSION 4

typedef struct pt_ctx pt_ctx;

typedef struct pt_opts {
    int device;                 /* HIP device ordinal */
    uint32_t blocks_per_cu;     /* persistent kernels' blocks per compute unit (0: library default) */
    const char* cache_dir;      /* code-object cache (NULL: $PT_KCACHE, else kcache/ beside the library) */
    uint32_t straggler_rounds;  /* rounds a sample runs beyond maxbounce + 1 for paths a hit from behind held up
                                   (0: PT_STRAGGLER_ROUNDS; PT_STRAGGLER_NONE: none, the flush closes them) */
    uint32_t flags;             /* PT_OPT_* */
    uint32_t reserved[2];
} pt_opts;
#define PT_STRAGGLER_NONE 0xFFFFFFFFu
/* Launch a sample's kernels one by one on the stream instead of replaying the captured hipGraph (for tools that
 * follow kernel launches but not graph launches). Same kernels, same order, same results. */
#define PT_OPT_DIRECT_LAUNCH 1u

/* A scene as the HIP emitter (shadytracer_amd/js/hipgen.js) produces it: HIP C++ defining sceneDist and
 * sceneFull, the parameter block they read, the material table and the light faces. */
typedef struct pt_scene_desc {
    const char* hip_src;
    const float* params;
    uint32_t nparams;
    const PtMaterial* materials;
    uint32_t nmaterials;
    const PtLightFace* faces;
    uint32_t nfaces;
    uint32_t flags;             /* PT_SCENE_* */
} pt_scene_desc;
/* The new scene draws the same picture as the one loaded (the reference's setScene({keepPicture}): a scene baked with
 * its values as constants): the accumulation is kept, not cleared. */
#define PT_SCENE_KEEP_PICTURE 1u

/* ABI version of this library (PT_ABI_VERSION). */
int pt_abi_version(void);

/* sizeof of the ABI's structures, for bindings to check their mirrors: 0 PtUniforms, 1 PtMaterial,
 * 2 PtLightFace, 3 PtCamera, 4 PtEnv, 5 PtHaze, 6 pt_scene_desc, 7 pt_opts, 8 PtGuideConfig, 9 PtOpticsScene,
 * 10 PtOpticsResult, 11 PtTuning, 12 PtOpticsSource, 13 PtOpticsTrace, 14 PtOpticsTally, 15 PtDevelop,
 * 16 PtDevelopInfo; 0 for anything else.
 * No counterpart in the reference (its uniform packing lives in the absent tinywebgpu submodule). */
size_t pt_struct_size(int which);

/* Creates an engine on a device and integrates its GGX albedo tables.
 * Replaces RAYTRACER(WG, _U, Q, scene) and R.init (lib/pathtracer/core/raytrace_main.js:121, :1291-1313). */
int pt_create(pt_ctx** out, const pt_opts* opts);

/* Frees everything the engine holds. Replaces R.releaseBuffers (raytrace_main.js:234-302). */
void pt_destroy(pt_ctx* ctx);

/* The last error's text, never NULL. Replaces the Error objects R.setScene / R.trace throw
 * (raytrace_main.js:1162-1176, :1579-1580). ctx may be NULL for errors of pt_create. */
const char* pt_last_error(const pt_ctx* ctx);

/* Compiles a scene's kernels for gfx950 into the code-object cache without a GPU, and writes the cached
 * file's path into path_out. The build step of _buildWavefrontPipelines (raytrace_main.js:585-658). */
int pt_compile_scene(const char* hip_src, const char* cache_dir, char* path_out, size_t path_cap);

/* Loads a scene: compiles (or finds in the cache) and loads its kernels, uploads its tables and clears the
 * accumulation (kept with PT_SCENE_KEEP_PICTURE). On a compile failure the previous scene stays loaded and
 * PT_ERR_COMPILE is returned (the reference's e.restored). Safe to call from a worker thread while the caller makes no
 * other call on the context. Replaces R.setScene(s, {keepPicture}) (raytrace_main.js:1145-1183). */
int pt_set_scene(pt_ctx* ctx, const pt_scene_desc* scene);

/* Updates the scene's parameter block in place (same length): the builder's per-node uniforms, without a
 * rebuild. Replaces the scene-uniform half of the uniform bag (raytrace_main.js:1997-2023). */
int pt_set_params(pt_ctx* ctx, const float* params, uint32_t nparams);

/* Compiles (or finds in the cache) the scene's kernels with its current parameter block folded in as constants and
 * runs those until the next pt_set_params or pt_set_scene, which go back to the kernels that read the block. Same
 * images (the same arithmetic on the same values); the distance function's parameter reads become immediates. No
 * reference counterpart (the reference keeps its scene parameters live uniforms): for a host about to render a
 * scene that will not move, after its edits. */
int pt_fold_params(pt_ctx* ctx);

/* Sets the engine uniforms (size must be sizeof(PtUniforms)); width/height resize the buffers.
 * Replaces the base half of the uniform bag and R.ensureBuffers (raytrace_main.js:234-302, :1997-2023). */
int pt_set_uniforms(pt_ctx* ctx, const PtUniforms* u, size_t size);

/* Clears the per-pixel accumulators and the sample counter. Replaces R.clearAccums (raytrace_main.js:1131). */
int pt_clear_accum(pt_ctx* ctx);

/* Enqueues `samples` samples (one camera ray per pixel each) on the engine's stream; returns at once.
 * Replaces R.trace(Q, uniforms, count) (raytrace_main.js:1574-2342). */
int pt_trace(pt_ctx* ctx, uint32_t samples);

/* Waits for everything enqueued. Replaces the drain on queue.onSubmittedWorkDone (lib/pathtracer/index.js:1005-1007). */
int pt_sync(pt_ctx* ctx);

/* Copies the FP64 accumulators: accum 4 doubles per pixel (mean r, g, b, samples), moments 2 per pixel (V, S
 * of the luminance; standard error sqrt(S V / (1 - S))). Either pointer may be NULL. npx must be width*height.
 * Replaces the readbacks of R.imageStats / R.readRegion (raytrace_main.js:876-956, :1417-1448). */
int pt_read_accum(pt_ctx* ctx, double* accum, double* moments, size_t npx);

/* Copies the engine's counters (PT_S_* indices, 64-bit each) since the last pt_reset_stats; n <= PT_S_WORDS.
 * Replaces the counter readback and wavefront log (raytrace_main.js:106-116, :2335-2341). */
int pt_read_stats(pt_ctx* ctx, uint64_t* stats, size_t n);
int pt_reset_stats(pt_ctx* ctx);

/* Runs `samples` samples kernel by kernel with HIP events around each launch on the engine's stream and
 * returns the summed milliseconds per kernel: ms[0] genrays, [1] march, [2] shade, [3] occlude, [4] resolve,
 * [5] the queues' binning, [6] the photon pass's map clear and emission, [7] surface, [8] the guide network's
 * inference and training, and of those [9] inference, [10] the training's gradient pass (forward and backward),
 * [11] the rest of training: collection, Adam, commit (n >= 5; the slots past n are dropped).
 * Replaces R.timeKernels (raytrace_main.js:549-557). */
int pt_time_kernels(pt_ctx* ctx, uint32_t samples, double* ms, size_t n);

/* Runs one sample's camera rays to their first event and reports, per ray, where each stopped (pos4: x, y, z,
 * optical length) and what it hit (info: state | material << 8, then the primitive; 0xFFFFFFFF none). n must be
 * the camera rays of a sample (width*height). The sample counter is not advanced. Replaces R.pick
 * (lib/pathtracer/core/pick.js) for every pixel at once. */
int pt_pick_rays(pt_ctx* ctx, float* pos4, uint32_t* info, size_t n);

/* ---- launch choices ----
 * No reference counterpart: the reference's dispatch sizes are fixed in its pipelines (wavefront.js workgroup sizes,
 * raytrace_main.js's 32-step passes). Here the persistent kernels' grid and their refill thresholds are chosen per
 * scene and image by measurement. */

/* The choices in force (PtTuning; size = sizeof(PtTuning)). */
int pt_get_tuning(pt_ctx* ctx, PtTuning* out, size_t size);
/* Sets them (blocksPerCu 1..32, refills 1..64); the next trace captures its graph anew. A uniforms' marchRefill of
 * 0 takes the refills from here. */
int pt_set_tuning(pt_ctx* ctx, const PtTuning* tuning, size_t size);
/* The autotuner: for the scene and uniforms set, times `samples` (1..64) samples of each candidate setting -- blocks
 * per CU in {2, 4, 6, 8, 12, 16}, then the march refill, then the occlusion refill, each in {1, 4, 8, 16, 24, 32, 48},
 * coordinate by coordinate from the setting in force -- keeps the fastest, and clears the accumulators (the timed
 * samples went into them; with guiding on, its training goes on through them). out (NULL or a PtTuning) receives the
 * choice, the candidates timed and the times before and after. */
int pt_autotune(pt_ctx* ctx, uint32_t samples, PtTuning* out, size_t size);

/* ---- path guiding (no counterpart in the reference: its guiding was removed, lib/pathtracer/README.md:853-861;
 * the engine's uniforms switch it on: `guide`, and `guideFrac` the share of guided directions) ---- */

/* Sets the guide network's training (PtGuideConfig; size = sizeof(PtGuideConfig)) and starts it afresh: weights
 * drawn from cfg->seed (He-uniform, biases 0), Adam's moments and step count zeroed. batch: a multiple of 64 in
 * [64, 2^20]; steps in [1, 64]; lr > 0; the box not empty. */
int pt_guide_config(pt_ctx* ctx, const PtGuideConfig* cfg, size_t size);
/* The network's FP32 master weights, 16169 floats in the order W1[80][64], b1, W2[64][64], b2, W3, b3, W4[64][41],
 * b4: read (pt_guide_weights) or replaced, with Adam's state reset (pt_guide_load). */
int pt_guide_weights(pt_ctx* ctx, float* theta, size_t n);
int pt_guide_load(pt_ctx* ctx, const float* theta, size_t n);
/* The network's forward pass on given vertices: raw 10 floats a row (position, view direction, normal,
 * roughness) in, 41 head outputs a row out (mixture logits, 8 x 3 mean vectors, 8 log-concentrations, the guided
 * fraction's logit). */
int pt_guide_forward(pt_ctx* ctx, const float* raw, size_t n, float* out41);
/* One Adam step on given training samples, 16 floats a row: position mapped into the box (3), roughness, view
 * direction (3), weight F, normal (3), 0, sampled direction (3), the BSDF's density there. n in [1, batch]. grad
 * (NULL or 16169 floats) receives the summed gradient the step applied; loss (NULL or 1 double) the batch's loss. */
int pt_guide_train(pt_ctx* ctx, const float* samples, size_t n, float* grad, double* loss);
/* The training's state after the last sample: out[0] Adam steps taken, out[1] training samples the sample
 * collected, out[2 + j] the loss of its step j (j < steps); n >= 2. */
int pt_guide_status(pt_ctx* ctx, double* out, size_t n);

/* ---- the developed picture ---- */

/* The accumulated picture developed to 8-bit RGBA (npx = width x height pixels, rows top first with d->flipY): with
 * d->autoExposure the exposure is metered from the picture (a log2-luminance histogram's trimmed average, highlights
 * protected, eased from the last call; the first call after pt_clear_accum sets it), else it is d->exposureEV. Then the
 * darkroom's stages that d->lightOn names (white balance and gains, the tone equaliser, Oklab saturation, vibrance and
 * hue, the vignette), the gamut clip, the tone curve, the output transfer, Levels, the curves (d->curveLut), the colour
 * wheels as an ASC CDL, the display's contrast, saturation and brightness, grain and dither; or, with d->measureMode,
 * the false-colour view of the illuminance or luminance. lightOn 0 is the display's own picture. info (NULL allowed)
 * says what was used and metered.
 * Replaces the develop pass and the auto-exposure kernels (lib/pathtracer/core/rendertoscreen.js displayPixel
 * :209-270, developPixel :423-455 with its stages :286-421 and tone curves :35-150; core/autoexposure.js
 * prog_ae_histogram / prog_ae_update :54-153; the settings from darkroom/grade.js lightUniforms :231-271 and
 * darkroom/curves.js bakeLut :75-87) for a picture that is saved rather than shown (darkroom/export/png.js takes its
 * result). */
int pt_develop(pt_ctx* ctx, const PtDevelop* d, size_t dsize, uint8_t* rgba, size_t npx, PtDevelopInfo* info);

/* pt_develop to 16 bits a channel: rgba16 holds npx RGBA pixels of uint16 (alpha 65535), round(v 65535) of what the
 * stages computed, with d's dither off (16 bits have no 8-bit steps to break).
 * Replaces the develop pass drawn into a float target for a 16-bit PNG (lib/pathtracer/develop.js developedRows and
 * developUniforms' `float` :73-77; darkroom/export/run.js :30-41 rounding to 16 bits). */
int pt_develop16(pt_ctx* ctx, const PtDevelop* d, size_t dsize, uint16_t* rgba16, size_t npx, PtDevelopInfo* info);

/* A view of the developed picture, outW x outH 8-bit RGBA (npx = outW x outH, rows top first with d->flipY):
 * view = {cx, cy, sx, sy}, the picture point (in picture pixels, y up) at the output's centre and the picture pixels
 * one output pixel spans. Each output pixel averages ceil(max(sx, sy)) squared taps (1..8 a side) over its
 * footprint, the backdrop grey 0.09 outside the picture; clipWarn stripes what clips (red at white, blue at black).
 * {w/2, h/2, w/outW, h/outH} is the whole picture, a thumbnail at a smaller output.
 * Replaces the develop pass's view (lib/pathtracer/core/rendertoscreen.js develop_prog :555-593; develop.js boxForSpan
 * :61 and viewFill :67; darkroom/view.js toUniform :73). */
int pt_develop_view(pt_ctx* ctx, const PtDevelop* d, size_t dsize, const float* view, uint32_t outW, uint32_t outH, uint32_t clipWarn,
                    uint8_t* rgba, size_t npx, PtDevelopInfo* info);

/* What pt_develop would write, counted: every `step`-th pixel each way developed as pt_develop does it (metering
 * included), hist[0..255] red, [256..511] green, [512..767] blue, [768..1023] the Rec.709 luma of the display values,
 * bin round(v 255); nbins = 4 x PT_HIST_BINS. A false-colour view (d->measureMode) is refused.
 * Replaces the darkroom's histogram pass (lib/pathtracer/core/rendertoscreen.js develop_histogram_prog :597-643;
 * develop.js histogramStep :97). */
int pt_develop_histogram(pt_ctx* ctx, const PtDevelop* d, size_t dsize, uint32_t step, uint32_t* hist, size_t nbins,
                         PtDevelopInfo* info);

/* ---- the optics trace over the loaded scene ---- */

/* Traces t->rays rays from the sources through the loaded scene -- the engine's own binning, march, surface kernels and
 * materials, with its uniforms in force (tolerance, maxiter, haze) -- and returns where their flux went: escaped, lost
 * (past the bounce cap, maxDist or maxiter), detected, absorbed in the haze, and per material the flux arriving and
 * leaving and the interactions (PtOpticsTally, lm). far (NULL when t->farMode is 0, else nfar = nG x nC doubles,
 * gamma-major) receives the far field, lm per (gamma, C) bin of the type C grid; hits (NULL, or hitCap x
 * PT_OPT_HIT_WORDS floats) the first min(t->detCap, hitCap) detector hits. The same seed traces the same rays. Sources
 * with no flux are left out; a trace needs one with flux.
 * Replaces createOptics(RT, WG) with setSources, run, results and readHits (lib/pathtracer/optics/index.js:92-825,
 * :434-477 run, :613-631 readHits, :687-697 setSources) over optics/kernels.js prog_opt_gen :504-531 (sources:
 * kernels.js SOURCE_RAY :270-420, optics/sources.js :1-60) and prog_opt_shade :533-723 (tallies, far field, detector
 * hits), and the far-field binning of optics/farfield.js:44-56. */
int pt_optics_run(pt_ctx* ctx, const PtOpticsSource* src, uint32_t nsrc, const PtOpticsTrace* t, size_t tsize,
                  PtOpticsTally* out, size_t osize, double* far, size_t nfar, float* hits, uint32_t hitCap);

/* ---- the optics bench ---- */

/* Traces n luminaire scenes (PtOpticsScene) in one persistent launch on `device` and returns each one's numbers
 * (PtOpticsResult): the flux balance, the far field's maximum, beam and field angles in the C0 and C90 planes, the
 * escaped light's chromaticity, CCT and Duv, the lux map's maximum and centre. far (NULL, or n x PT_OPT_NG x
 * PT_OPT_NC doubles) receives each scene's far field, lm per (gamma, C) bin; lux (NULL, or n x PT_OPT_LUX^2) its lux
 * map at z = 1 m, lx; ms (NULL, or 2 doubles) the walk kernel's time and the numbers kernel's, in ms (HIP events on
 * the launch's stream; uploads and allocation outside). Needs no engine. On failure pt_optics_error() says why.
 * Replaces createOptics's run and results (lib/pathtracer/optics/index.js) over optics/kernels.js:43-87 (gen, shade,
 * tallies), optics/farfield.js:33-168 (farBin, intensity, farStats), optics/detmap.js:104-213 (the map) and
 * core/colorimetry.js:58-104 (xyToCctDuv). */
int pt_optics_trace(int device, const PtOpticsScene* scenes, uint32_t n, PtOpticsResult* out, double* far, double* lux,
                    double* ms);
const char* pt_optics_error(void);

/* The same as pt_optics_trace -- the same kernels, the same numbers -- with no allocation of any kind: every device
 * buffer lives in one array of PT_OPT_STATIC_BYTES that the library holds on each GPU it runs on until the process
 * ends (the last call's contents stay in it; what a call sums into it clears first). No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Limits: at most PT_OPT_STATIC_SCENES scenes a call, and their photons rounded up to chunks of 4096
 * at most PT_OPT_STATIC_CHUNKS chunks in all (2^27 photons); beyond them PT_ERR_ARG with a message, as for any bad
 * argument. No timings (ms): events would be an allocation. Calls on one GPU share its array: this library serialises
 * its own callers in one process; overlapping calls through any other route race. */
#define PT_OPT_STATIC_SCENES 8u
#define PT_OPT_STATIC_CHUNKS 32768u
#define PT_OPT_STATIC_BYTES (11u << 20)
int pt_optics_trace_static(int device, const PtOpticsScene* scenes, uint32_t n, PtOpticsResult* out, double* far, double* lux);

#ifdef __cplusplus
}
#endif
#endif
