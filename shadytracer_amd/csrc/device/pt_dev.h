// The device pointer table: everything a kernel needs, by one pointer (a kernel argument, read through scalar
// loads). Internal to the library; the host fills it (pt_api.cpp).
#pragma once

struct PtDev {
    PtUniforms u;
    unsigned int tilesX, tilesY;
    float invTX, invTY;
    unsigned int nrays;         // camera rays per sample
    unsigned int capacity;      // path slots
    unsigned int nmats, nfaces;
    float faceTotal;
    unsigned int gridBlocks;    // persistent kernels' grid
    unsigned int refillMarch;   // march: refill a wave once this many lanes are idle
    unsigned int refillOcclude; // occlude: the same
    unsigned int mfReady;
    unsigned int pad1;
    // path state, structure of arrays indexed by slot
    float* pos;        // 4/slot: xyz, ltot (optical length so far)
    float* dir;        // 4/slot: xyz, n (index of the medium the ray is in)
    float* beta;       // 4/slot: rgb, prevPdf
    float* rad;        // 4/slot: rgb, vtxLtot
    unsigned int* meta;// 4/slot: pix, counts (steps | diffuse << 16), bounces | state << 8 | flags << 16, tag
    float* aux;        // 4/slot: sunmis, tmax (haze free flight), lambda, spare
    // queues of slots
    unsigned int* qMarch0;
    unsigned int* qMarch1;
    unsigned int* qShadow;
    unsigned int* qSorted;          // this round's march queue, binned by the rays' node masks
    unsigned long long* maskBuf;    // per slot: the nodes its current segment can reach
    float* hitRec;                  // per slot, 4 words: the surface kernel's normal and material word
    unsigned int* binHist;          // PT_BINS counts, then (scan) the bins' write offsets
    unsigned int* binOff;
    // shadow rays by slot
    float* shOrig;     // 4/slot: xyz, tmax
    float* shDir;      // 4/slot: xyz, n
    float* shRad;      // 4/slot: rgb contribution, pix (bits)
    unsigned int* counters;
    unsigned long long* stats;
    // per-sample frame and FP64 accumulators
    float* frame;      // 4/px: rgb sum, count (u32 bits)
    double* accum;     // 4/px: mean rgb, count
    double* moments;   // 2/px: V, S
    // scene
    const float* params;
    const PtMaterial* mats;
    const PtLightFace* faces;
    const float* mf;   // GGX albedo tables
    unsigned int* seedCur;  // 4 words: this sample's seed
    // SPPM photon map
    unsigned int* sppmHeads;        // 3 x PT_SPPM_HEADS: the cells' list heads (record index + 1; 0 empty), the cells'
                                    // record counts, the sub-lists' heads
    float* sppmRecs;                // 8 words a photon: position, power rgb, direction to the light (oct 2 x 16),
                                    // normal (oct 2 x 13) << 5 | sub-list
    unsigned int* sppmNext;         // 2 words a photon: its successor in its cell's list and in its sub-list
    unsigned int sppmCap;           // records the map holds
    unsigned int nphotons;          // photons emitted in a pass (slots [0, nphotons) of the path state)
    // path guiding (pt_guide.h)
    PtGuideConfig g;
    float* gMaster;        // the network's FP32 master weights (16104)
    float* gM;             // Adam's moments
    float* gV;
    float* gPart;          // PT_G_PARTS rows of partial gradients
    float* gGradOut;       // NULL, or 16104: the last step's summed gradient
    unsigned short* gW;    // the BF16 copy, each layer transposed, for the MFMA
    float* gBias;          // 256: b1, b2, b3, b4 (padded)
    float* gMix;           // per slot, 11 float4: the vertex's mixture (pt_guide.h guideStoreMix)
    float* gRec;           // recorded vertices: [path][vertex], 7 float4
    float* gPath;          // per recorded path, 2 float4: its total (rgb, done), its light sampling's (rgb)
    unsigned int* gVerts;  // per recorded path: its vertices recorded
    float* gBatch;         // training samples, 4 float4 each
    double* gSumF;         // per step: the sum of the training weights of its batch
    double* gLoss;         // per step: the batch's loss
    unsigned int gStride, gPaths, gMaxV, gBatchCap;
};

// The optics trace's own table (pt_optrace.h): the sources, the trace's settings and its tallies, for one batch.
struct PtOptDev {
    const PtOpticsSource* src;      // sources with flux, each's cumulative share of the total in cdf
    const float* cdf;
    unsigned int nsrc;
    unsigned int rays;              // this batch's rays: slots [0, rays)
    unsigned int firstId;           // the batch's first ray id (a ray's draws are keyed by its id)
    unsigned int bounces;
    unsigned int seed[4];
    PtOpticsTrace t;                // the settings, axis and C0 normalised
    unsigned long long* tally;      // PT_OT_WORDS
    unsigned long long* far;        // nG x nC, gamma-major
    float* hits;                    // detCap x PT_OPT_HIT_WORDS (flux in units of ray weight)
};

