// Plain-old-data structures of the C ABI (include/pt.h), shared by the host library and the device code that
// hipRTC compiles at scene load. Only fixed-width scalars and explicit padding: the same bytes on both sides.
//
// Conventions (the reference's): world Z-up, metres, linear Rec.709 RGB whose Y is cd/m^2
// (lib/pathtracer/README.md:237-238, :290-309; core/sdf.js:619-621).
#ifndef PT_TYPES_H
#define PT_TYPES_H

#define PT_MAX_MATERIALS 64
#define PT_MAX_FACES 64
#define PT_MF_NMU 32
#define PT_MF_NR 32
#define PT_MF_NIOR 16
// Ec[ior][r][mu], EcAvg[ior][r], E1[r][mu]
#define PT_MF_EC 0
#define PT_MF_ECAVG (PT_MF_NIOR * PT_MF_NR * PT_MF_NMU)
#define PT_MF_E1 (PT_MF_ECAVG + PT_MF_NIOR * PT_MF_NR)
#define PT_MF_WORDS (PT_MF_E1 + PT_MF_NR * PT_MF_NMU)

// Path states (the reference's RAY_* of core/tracebasics.js, renumbered)
#define PT_ACTIVE 0u
#define PT_ESCAPED 1u
#define PT_MAXITER 2u
#define PT_TERMINATED 3u
#define PT_SCATTER 4u   // a collision in the haze box, to be shaded as a medium vertex

// Device counters (u32 words)
#define PT_C_MARCH0 0
#define PT_C_MARCH1 1
#define PT_C_SHADOW 4
#define PT_C_SAMPLE 6      // samples resolved since the accumulators were cleared (drives the per-sample seed)
#define PT_C_PHOTON 7      // photons stored in the SPPM map this pass
#define PT_C_RADIUS 8      // this sample's SPPM gather radius r_k (float bits), set by the photon pass
#define PT_C_GUIDE_N 9     // guide training samples collected this sample
#define PT_C_GUIDE_T 10    // Adam steps the guide network has taken
#define PT_SPPM_HEADS (1u << 20)   // the photon map's hash-grid lists (the reference's SPPM_HEADS, sppm.js:79)
#define PT_BINS 1024          // bins of the march queue's counting sort by node mask
#define PT_BIN_BLOCKS 256     // blocks of that sort (each sorts one contiguous chunk of the queue)
#define PT_PARTS 64           // partitions of a queue that the persistent kernels' waves take batches from
#define PT_C_HEADS_MARCH 64   // PT_PARTS fetch heads of the march queue
#define PT_C_HEADS_SHADOW 128 // PT_PARTS fetch heads of the shadow queue
#define PT_C_WORDS 256

// 64-bit statistics
#define PT_S_SEGMENTS 0      // path segments: march events handed to shade (hit, escape, cap, haze collision)
#define PT_S_SHADOWSEGS 1    // shadow-ray segments traced
#define PT_S_STEPS 2         // sphere-tracing steps of path rays
#define PT_S_SHADOWSTEPS 3   // sphere-tracing steps of shadow rays
#define PT_S_NONFINITE 4     // deposits refused as NaN/Inf/negative luminance
#define PT_S_MAXITER 5       // rays that ran out of steps or length
#define PT_S_PATHS 6         // camera paths completed (deposited)
#define PT_S_LEAVING 7       // hits from behind a surface, moved on without a bounce
#define PT_S_GATHERS 8       // SPPM camera gathers
#define PT_S_PHOTONS 9       // SPPM photons stored
#define PT_S_FLOPS 10        // FP32 operations of the path rays' sphere tracing (scene SDF and step arithmetic)
#define PT_S_SHADOWFLOPS 11  // the same for shadow rays
#define PT_S_PHOTONSEGS 12   // photon segments traced (SPPM photon passes)
#define PT_S_GATHERCUT 13    // SPPM gathers that reached the bound on records visited
#define PT_S_RADIUS 14       // the last photon pass's gather radius r_k (float bits; set, not summed)
#define PT_S_OVERFLOW 15     // SPPM stores dropped for want of room (the map holds 2 records a photon)
#define PT_S_GSTEPS 16       // guide network: Adam steps taken
#define PT_S_GUIDED 17       // path vertices sampled with guiding on
#define PT_S_GRECORDS 18     // path vertices recorded for training
#define PT_S_FLUSHED 19      // camera paths still going when a sample's rounds ran out, deposited as they stood
#define PT_S_SHADOWCUT 20    // shadow rays cut short by maxiter, counted as reaching their light (wavefront.js:1640-1645)
#define PT_S_GINFER 21       // guide network: path vertices run through the network for their mixture (inference)
#define PT_S_GTILES 22       // guide network: tiles of 64 rows the inference ran (the rows the matrix cores computed)
#define PT_S_GTRAIN 23       // guide network: training rows through the forward and backward pass (all Adam steps)
#define PT_S_WORDS 24
// Rounds of march -> shade a sample runs beyond maxbounce + 1: a hit from behind moves a path on without a bounce,
// so a path that met some still finishes; one still going after them is flushed (PT_S_FLUSHED)
#define PT_STRAGGLER_ROUNDS 1

typedef struct PtCamera {
    float pos[3];
    float fx;            // invProjection[0][0]: aspect * tan(fov/2)
    float right[3];
    float fy;            // invProjection[1][1]: tan(fov/2)
    float up[3];
    unsigned int type;   // 0 perspective, 1 orthographic
    float fwd[3];        // from the target towards the camera; the view direction is -fwd
    float lensR;         // the thin lens's radius, m (0: a pinhole; perspective only)
    float focus;         // the distance in focus along the view direction, m
    float pad[3];
} PtCamera;

// A material as the shade kernel reads it (the reference's Material2 subset, core/materials.js:12-110).
typedef struct PtMaterial {
    float base[3];        // baseColor
    float metallic;
    float roughness;
    float transmission;
    float ior;
    float specular;       // 0.5 is the dielectric default (F0 = 0.04 at ior 1.5)
    float emissive[3];    // radiance, cd/m^2 as Y (colour scaled to emissiveLuminance)
    float abbe;           // > 0: dispersive (Cauchy through n_d and V_d)
    int face;             // light face this emissive material belongs to, -1 none
    float specularTint;
    float attenuation;    // > 0: absorbs inside, the depth (metres) at which it has filtered light to base (Beer-Lambert)
    float pad[1];
} PtMaterial;

// A light's face for next-event estimation (the reference's LightFace, core/tracebasics.js:266-277).
typedef struct PtLightFace {
    float centre[3];
    float halfx;
    float u[3];
    float halfy;
    float v[3];
    unsigned int shape;   // 1 disc (halfx = radius), 2 rectangle
    float n[3];
    unsigned int twoSided;
    float le[3];
    float weight;
} PtLightFace;

// Environment: 0 black, 1 uniform sphere, 2 uniform upper hemisphere (z > 0), 3 the studio -- L (1 + top d.z), never
// below 0, over the whole sphere (core/environment.js studioLight:165-174) -- 4 the sun & sky (environment.js
// SUNSKY_FN:110: L is the sunlight above the air, lux; the sky scattered once in sixteen bands) -- plus an optional
// sun.
typedef struct PtEnv {
    unsigned int type;
    float L;              // luminance, cd/m^2
    float col[3];         // colour with luma 1
    float sunDir[3];      // towards the sun
    float sunCos;         // cos of the sun's angular radius
    float sunE;           // illuminance on a surface facing the sun, lux (0: no sun)
    float sunRadius;      // the beam's radius (0: no aperture, the sun lights everything)
    float sunPoint[3];    // the centre of the beam's disc: with an aperture, the sun is seen along the rays that
                          // cross the disc of radius sunRadius about sunPoint, perpendicular to sunDir (and the
                          // photon pass emits the beam from that disc)
    float top;            // the studio's top light, 0..1 (type 3)
    float pad;
    float sunL[3];        // type 4: the sun disc's radiance, RGB cd/m^2 (other types: sunE over the solid angle, white)
    float sunSin2;        // sin^2 of the sun's angular radius: a direction d is in the disc when |d x sunDir|^2 <= this
    float sunOmega;       // the disc's solid angle, 2 pi (1 - cos), worked out in double
    float skyAerosol;     // type 4: the haze's optical depth at 550 nm
    float skyMs;          // type 4: the air mass the sunlight crosses
    float skyUp;          // type 4: 1 while the sun is up, fading to 0 over its last 3 degrees below the horizon
} PtEnv;

// The homogeneous haze box (the reference's haze.js hazeUniforms): axis-aligned, centre and half size.
typedef struct PtHaze {
    float centre[3];
    float sigmaS;         // scattering coefficient, 1/m (0: no haze)
    float half[3];
    float g;              // Henyey-Greenstein asymmetry
    float albedo;
    float pad[3];
} PtHaze;

// The persistent kernels' launch choices: what pt_autotune picks for the scene and image in hand, or a host sets with
// pt_set_tuning. A uniforms' marchRefill of 0 takes the refills from here; a nonzero one overrides both.
typedef struct PtTuning {
    unsigned int blocksPerCu;     // blocks of 256 threads a compute unit for march, surface, shade, occlude (1..32)
    unsigned int refillMarch;     // a march wave takes new rays once this many of its 64 lanes are idle (1..64)
    unsigned int refillOcclude;   // the same for the occlusion kernel
    unsigned int candidates;      // pt_autotune: the settings it timed
    double msPerSample;           // pt_autotune: the chosen setting's time a sample (ms)
    double msBefore;              // pt_autotune: the setting in force before it, timed alike
} PtTuning;

// The engine's uniforms: the subset of the reference's schema (core/uniforms.schema.js:37-418) this path reads.
// The path-guiding network's training (set with pt_guide_config): the box vertex positions are encoded in, Adam's
// learning rate, samples a step and steps a sample, which paths record their vertices, the warm-up before guided
// sampling starts and the weights' initialisation seed.
typedef struct PtGuideConfig {
    float boxMin[3];
    float lr;              // 1e-2
    float boxMax[3];
    unsigned int batch;    // training samples an Adam step (2^16)
    unsigned int steps;    // Adam steps a sample (each on its own batch)
    unsigned int stride;   // one camera path in `stride` records (0: chosen from batch x steps and the image)
    unsigned int warmup;   // Adam steps before guided sampling starts
    unsigned int seed;     // weights' initialisation
} PtGuideConfig;

// One luminaire of the optics bench (metres, lumens; the LED at the origin emitting along +z): a Lambertian square
// LED, a lens of two conic faces (z = vertex + sag(r), sag = c r^2 / (1 + sqrt(1 - (1 + k) c^2 r^2)); c 0 is a
// plane) with an absorbing rim, an optional paraboloid reflector (focus at the LED) and an absorbing board at z = 0.
typedef struct PtOpticsScene {
    float ledHalf;         // half the LED's side (0.5 mm)
    float flux;            // lm
    float xyz[3];          // the LED's colour as X, Y, Z of one lumen (Y = 1)
    float lensZ;           // the front face's vertex
    float lensT;           // centre thickness (front vertex to back vertex)
    float lensD;           // diameter
    float c1, k1;          // front face: curvature (1 / R), conic constant
    float c2, k2;          // back face
    float ior;
    float reflF;           // paraboloid focal length (0: no reflector)
    float reflR0;          // its base radius (the hole round the LED)
    float reflZ1;          // its mouth's z
    float rho[3];          // its reflectance per X, Y, Z channel
    float boardR;          // the absorbing board's radius at z = 0
    unsigned int photons;
    unsigned int seed;
} PtOpticsScene;

// An optics bench result (FP64 on the GPU): flux balance, the far field's numbers (farfield.js farStats) and colour
// (colorimetry.js xyToCctDuv), the lux map's.
typedef struct PtOpticsResult {
    double emitted, escaped, absorbed, efficiency;   // lm, lm, lm, escaped / emitted
    double Imax, gammaMax, cMax;                     // cd at the brightest bin, and where (deg)
    double beam[2], field[2];                        // full widths at 50 % and 10 % of Imax in the C0-C180 and
                                                     // C90-C270 planes, deg (-1: the plane never falls that far)
    double X, Y, Z, x, y, cct, duv;                  // the escaped light's tristimulus totals and chromaticity
    double luxMax, luxCentre;                        // the map at 1 m: its largest cell, the mean of its 4 central
    double photons;
    double events;                                   // surface tests the walks made (one a bounce: every surface tried)
    double flops;                                    // FP32 operations the walks performed, from per-step source counts
                                                     // (0 from the CPU oracle)
} PtOpticsResult;
#define PT_OPT_NG 181        // far field gamma rows (0, 1, ..., 180 deg; the first and last half wide)
#define PT_OPT_NC 360        // far field C columns (0, 1, ..., 359 deg)
#define PT_OPT_LUX 256       // lux map cells a side over [-1, 1] m at z = 1 m

// ---- The optics trace over the scene (optics/index.js createOptics; optics/kernels.js prog_opt_gen/prog_opt_shade):
// rays from sources walk the loaded scene through the engine's own march, surface and materials; where their flux goes
// is tallied per material, into the far field and at detectors.
// A source (optics/sources.js defaultSource / sourceRecords): a shape at a pose, an angular distribution about its axis,
// a flux. Every ray carries the same share of the total flux, so the distribution is in where rays go.
typedef struct PtOpticsSource {
    float pos[3];
    unsigned int shape;          // 0 point, 1 disc (size[0] its radius), 2 rectangle (size: half sizes along u, axis x u)
    float axis[3];               // the emission axis (normalised by the library)
    unsigned int distribution;   // 0 collimated (along the axis), 1 uniform in solid angle within the cone,
                                 // 2 density ~ cos^n within the cone (n 1 over 90 degrees: Lambertian)
    float u[3];                  // a tangent (any vector not along the axis)
    float halfAngle;             // the cone's half angle, degrees, 0..90
    float size[2];
    float n;                     // the cosine power (distribution 2)
    float flux;                  // lm (a source of no flux is left out)
    float colour[3];             // the light's hue, linear Rec.709 (scaled to luminance 1 by the library)
    float emitN;                 // the index of the medium it emits into (>= 1)
    unsigned int twoSided;       // both faces emit, half the rays each
    unsigned int onSurface;      // the source is a sheet of the scene: rays start two march tolerances off it
    unsigned int pad[2];
} PtOpticsSource;
// What a trace is asked (optics/index.js O.settings).
typedef struct PtOpticsTrace {
    unsigned int rays;           // rays traced in all (the library runs them in batches)
    unsigned int bounces;        // interactions a ray is followed through (64); one past them it is LOST
    unsigned int seed;
    unsigned int detCap;         // detector hits recorded at most (more are counted)
    float maxDist;               // the longest optical path, metres; a hit beyond it is LOST (0: no limit)
    unsigned int farMode;        // far field: 0 off, 1 every escaping ray, 2 every ray arriving at detector material
                                 // farMaterial, by the direction it arrives along
    unsigned int farMaterial;
    unsigned int nC, nG;         // C bins about the axis (centred on 0, 360/nC, ...); gamma bins centred on
                                 // 0, 180/(nG-1), ..., 180 degrees, the two end ones half as wide
    float farAxis[3];            // gamma = 0 (normalised by the library)
    float farC0[3];              // C = 0 (made square to the axis by the library)
    unsigned int detectors[2];   // one bit per material id: a ray arriving at such a material from outside stops
                                 // there, all of its flux DETECTED and the hit recorded
    unsigned int marchSteps;     // sphere-tracing steps a ray takes over its whole path before it is LOST (the engine's
                                 // maxiter for the trace's rays; 0: 2000, optics/index.js settings.marchSteps)
    unsigned int pad;
} PtOpticsTrace;
// Where the flux went (optics/index.js readTallies / scaled), in lm: emitted = escaped + lost + volume + the sum over
// materials of (incident - leaving); absorbed = that sum less what the detectors caught, plus volume.
typedef struct PtOpticsTally {
    double emitted, escaped, lost, detected, volume, absorbed;
    unsigned long long rays;
    unsigned long long escapedRays, lostRays, interactions, absorbedRays, detectedRays;
    unsigned long long detectorHits;       // hits at detectors, recorded or not (recorded: min(this, detCap))
    unsigned long long pad;
    double incident[PT_MAX_MATERIALS];     // per material id: the flux arriving, and leaving after the interaction
    double leaving[PT_MAX_MATERIALS];
    unsigned long long hits[PT_MAX_MATERIALS];
} PtOpticsTally;
#define PT_OPT_HIT_WORDS 8   // a recorded detector hit: position xyz (m), flux (lm), the arriving direction xyz, material
// The device tallies: fixed point, PT_OPT_FX per unit of ray weight in 64-bit integer atomics (exact, and the same sum
// in any order). Word indices:
#define PT_OPT_FX 16777216.0
#define PT_OT_ESCAPED 0
#define PT_OT_LOST 1
#define PT_OT_DETECTED 2
#define PT_OT_VOLUME 3
#define PT_OT_CESC 4
#define PT_OT_CLOST 5
#define PT_OT_CINTER 6
#define PT_OT_CABS 7
#define PT_OT_CDET 8
#define PT_OT_HITS 9
#define PT_OT_MAT 16           // then per material: incident, leaving, hits
#define PT_OT_WORDS (PT_OT_MAT + 3 * PT_MAX_MATERIALS)

// ---- The developed picture (core/rendertoscreen.js displayPixel and developPixel, core/autoexposure.js,
// darkroom/grade.js lightUniforms, darkroom/curves.js bakeLut): exposure, the darkroom's stages, tone curve, 8-bit.
// lightOn's bits (grade.js LIGHT_*): which darkroom stage runs; a stage off is never computed, so with lightOn 0 the
// picture is the display's own.
#define PT_LIGHT_ANY 1u
#define PT_LIGHT_WB 2u
#define PT_LIGHT_EQ 4u
#define PT_LIGHT_LAB 8u
#define PT_LIGHT_LEVELS 16u
#define PT_LIGHT_CURVES 32u
#define PT_LIGHT_CDL 64u
#define PT_LIGHT_VIGNETTE 128u
#define PT_LIGHT_GRAIN 256u
#define PT_LIGHT_DITHER 512u
#define PT_LIGHT_CURVE_LUMA 1024u
#define PT_LIGHT_ALL 2047u
#define PT_CURVE_LUT 1024        // entries per curve of curveLut: master, red, green, blue
#define PT_HIST_BINS 256         // a developed picture's histogram: R, G, B and luma, this many bins each
typedef struct PtDevelop {
    unsigned int toneMap;        // 0 linear (clipped at white), 1 PBR Neutral, 2 AgX, 3 AgX Punchy, 4 ACES (fitted),
                                 // 5 Reinhard (luminance), 6 Hable, 7 Hable with tone7's white, toe and shoulder
                                 // (the reference's toneMap numbering)
    unsigned int autoExposure;   // 1: metered from the picture (the exposure eases by aeSpeed a call); 0: exposureEV
    float exposureEV;            // EV100 when not metered: white is 1.2 x 2^EV cd/m^2
    float aeSpeed;               // share of the gap to the metered EV closed a call (1: jump; the first call jumps)
    float aeHeadroom;            // stops the 95 % quantile may sit above white before the meter darkens the picture
    unsigned int lightOn;        // PT_LIGHT_* bits; 0: the display's picture, bit for bit
    unsigned int flipY;          // 1: rows top first (images, PNG); 0: as accumulated, bottom first
    unsigned int transfer;       // the output transfer: 0 sRGB, 1 gamma 2.2, 2 gamma 2.4, 3 linear
    float contrast, saturation, brightness;  // the display's grading on the encoded picture (0 each: none)
    unsigned int measureMode;    // 0 the picture; false colour of 1 the illuminance (lux), 2 the luminance (cd/m^2)
    float measureMin, measureMax;            // the false colour's log scale, its ends
    unsigned int measureBands;   // 0 continuous, else that many flat bands
    float wb[9];                 // white balance and gains: rows of the 3 x 3 applied to the exposed light
    float eq[8];                 // the tone equaliser's gains in stops at -7 .. 0 EV below white
    float lab[3];                // Oklab: saturation, vibrance, hue (radians)
    float tone7[3];              // curve 7: its white (linear), toe, shoulder
    float levels[5];             // Levels: input black, input white, 1 / gamma, output black, output white
    float cdlSlope[3], cdlOffset[3], cdlPower[3];  // the colour wheels as an ASC CDL
    float vignette[3];           // stops at the corners, where it starts (0..1 of the way out), round (0..1)
    float grain[2];              // amount, cell size in picture pixels (>= 1)
    float pad[7];
    float curveLut[4 * PT_CURVE_LUT];        // the curves baked: master, red, green, blue on the encoded value
} PtDevelop;
// What developing measured: the exposure used, the meter's trimmed log2 average and protected quantile, pixels metered.
typedef struct PtDevelopInfo {
    float ev, log2Average, log2High, metered;
} PtDevelopInfo;

typedef struct PtUniforms {
    unsigned int width, height;
    unsigned int maxbounce;     // 10
    unsigned int maxiter;       // 400 march steps per path
    float stepscale;            // 0.9
    float eps;                  // hit tolerance setting, metres (0: epsAuto)
    float epsAuto;
    float envRadius;            // metres
    unsigned int nee;           // 1
    unsigned int rr;            // 0
    unsigned int rrStart;       // 3
    unsigned int diffuseCap;    // 0: none
    float fireflyClamp;
    float fireflyWhite;
    unsigned int maxaverages;   // 0: unlimited
    unsigned int seed;
    PtCamera cam;
    PtEnv env;
    PtHaze haze;
    float faceShare;            // share of light sampling that goes to light faces (0.5)
    unsigned int sppm;          // 1: camera paths after delta-only chains gather from the photon map (the reference's);
                                // 2: every non-delta camera vertex gathers (caustics seen after a diffuse bounce too)
    float sppmRadius;           // r_1: the first sample's gather radius, metres
    unsigned int sppmPhotons;   // photons emitted in each sample's photon pass
    unsigned int marchRefill;   // lanes a wave lets go idle before it fetches more rays
    unsigned int guide;         // 1: MLP path guiding mixed with BSDF sampling
    float guideFrac;            // the largest share of guided directions (the network picks its own below it)
    float sppmAlpha;            // the radius schedule's alpha: r_k^2 = r_1^2 prod_{i<k} (i + alpha) / (i + 1)
    float relax;                // over-relaxed sphere tracing: steps of relax * stepscale * |d| while the free balls
                                // overlap (Keinert et al. 2014); 0 or 1 the plain march, at most 1.95
    unsigned int measureMode;   // the reference's measure views (uniforms.schema.js measureMode): 0 or 2 the light as it
                                // is; 1 the illuminance view -- the camera's first hit a white Lambertian for that
                                // bounce, so a pixel reads E / pi there, and a camera ray that meets nothing 0
    unsigned int lt;            // 1: light tracing (uniforms.schema.js lt): sun photons that passed a delta event connect to
                                // the camera where they first meet a surface or the haze; camera paths leave that light
                                // (first vertex, then delta events to the sun) to them. Pinhole and orthographic cameras
} PtUniforms;


#endif
