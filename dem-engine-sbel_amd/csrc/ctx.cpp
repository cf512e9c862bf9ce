// Host side of the C ABI declared in include/deme/deme.h: device memory ownership, stage orchestration, error mapping
// and per-kernel event timing. One context == one device + one HIP stream; no global mutable state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/deme/deme.h"
#include "deme_internal.h"
#include "jit/jit_gen.h"

using namespace deme_int;

namespace {

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t n, double grow = 1.25) {
        if (n <= cap && p) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            cap = 0;
        }
        size_t want = std::max<size_t>((size_t)(n * grow), 16);
        hipError_t e = hipMalloc((void**)&p, want * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        cap = want;
        return hipSuccess;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    // grows to hold n elements, keeping the first `keep` (synchronises the stream when it reallocates)
    hipError_t grow_keep(size_t n, size_t keep, hipStream_t st) {
        if (n <= cap && p) return hipSuccess;
        DBuf<T> nb;
        hipError_t e = nb.ensure(n);
        if (e != hipSuccess) return e;
        if (keep && p) {
            e = hipMemcpyAsync(nb.p, p, keep * sizeof(T), hipMemcpyDeviceToDevice, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) {
                nb.release();
                return e;
            }
        }
        release();
        p = nb.p;
        cap = nb.cap;
        return hipSuccess;
    }
};

struct TimedLaunch {
    int kid;
    hipEvent_t s, e;
};

float mean_radius(const std::vector<float>& r) {
    double sum = 0.0;
    for (float x : r) sum += x;
    return r.empty() ? 0.f : (float)(sum / (double)r.size());
}

}  // namespace

// Rounding allowance of the candidate lists' displacement test: positions are doubles near 1 m, so a nanometre lies
// far above their rounding and far below any skin.
constexpr double SKIN_SLACK = 1e-9;

struct deme_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool ownStream = false;
    std::string err;
    // the allocation-free scan's fixed workspace on this context's device (looked up at creation)
    uint64_t* noallocScratch = nullptr;
    uint64_t noallocElems = 0;

    deme_sim_params sp{};
    KParams kp{};
    bool haveParams = false, haveTemplates = false, haveState = false, haveMaterials = false, haveFamilies = false;
    bool maskAny = true;  // some family pair is masked off
    // the non-detection step replayed from a captured HIP graph (deme_set_step_graphs); any change of buffers,
    // kernel parameters or kernels bumps graphGen and the next such step captures again
    bool useGraphs = false;
    bool keepRecords = false;  // the last step of each deme_step call stores per-contact records
    uint64_t graphGen = 1, capturedGen = 0;
    hipGraphExec_t stepExec = nullptr;
    KParams capturedKp{};
    // A whole detection interval (the detection pass and the f steps up to the next one) captured in one graph per
    // list parity: kernels after the pass are guarded on its counters (grids cover the buffers' capacity), so nothing
    // needs the host between the pass and the steps. A replay whose pass failed or outgrew the buffers moves nothing;
    // the host then runs that interval directly (regrowing) and captures again.
    struct Cycle {
        hipGraphExec_t exec = nullptr;
        uint64_t gen = 0;
        int f = 0;
        KParams kp{};
        SpecGuard guard{};
        std::vector<const void*> ptrs;
    } cycle[2];
    uint64_t cycleReplays = 0, cycleFallbacks = 0, cycleCaptures = 0;
    // After a replay falls back, the next cycleSkip intervals run directly; the run doubles with each fallback in a
    // row (up to 64 intervals) and resets on a replay that completes.
    uint32_t cycleSkip = 0, cycleBackoff = 0;
    // Candidate-list detection: pairs within `skin` of touching, kept from a build until some sphere has moved half
    // the skin; each pass then tests only them. skinSet < 0: automatic (0.2 x the mean radius); 0: off (the cell
    // sweep every pass).
    // The automatic skin (any skin gives the same contacts) adapts at each build, within 0.05 to 0.8 x the mean
    // radius, to the lists just retired. Lists that die within a pass widen it 1.5x; lists that outlast 400 passes
    // narrow it 0.75x. Lists that last 30 passes or more leave it alone (their builds cost little). Shorter-lived
    // lists steer it by their cost per pass -- candidates times (1 + BUILD_PASSES / the passes they lasted), a build
    // costing about as much as BUILD_PASSES passes over its candidates: it moves 1.25x (widening first when lists
    // last fewer than 8 passes) while that cost falls by more than 2%, turns back when it rises by more than 2%, and
    // stops when it changes less.
    // With the automatic skin, lists that last fewer than 4 passes at the widest skin do not pay for their builds:
    // the cell sweep then runs every pass for cellPassesLeft passes (64, doubling while that repeats, up to 4096),
    // after which the lists are tried again.
    double skinSet = -1.0, skinAuto = 0.0;
    double skinCostPrev = -1.0;  // cost per pass of the lists before the ones just retired (-1: none)
    int skinDir = 0;
    uint32_t cellPassesLeft = 0, cellBackoff = 0;
    bool candValid = false, rebuildNext = true;
    uint64_t nCand = 0, candBuilds = 0, passesSinceBuild = 0;
    // passes the lists before the latest build lasted (~0 before the second build) and whether the pass in hand built
    uint64_t lastLifetime = ~0ull;
    bool builtThisPass = false;
    DBuf<CdRec> pos;
    // build positions in fixed point (coordinate = q * posUnit, posUnit a power of two set at each build)
    DBuf<BuildPos> pos0;
    double posUnit = 0.0;  // (from the bin grid: deme_set_params)
    // the common translation the displacement test subtracts: slot p (3 doubles) written by the pass at list parity p
    // (the midpoint of its displacement extremes since the build), read by the next pass; both zero after a build
    DBuf<double> dispRef;
    DBuf<uint2> meta;
    DBuf<uint32_t> candStart, candB;
    // The build's candidates: candA (the row's sphere), transposed (tC[tStart[B]..tStart[B + 1]]: indices of the
    // candidates with partner B, ascending A; tPos[c]: c's slot there). Per list parity the touching flags (candAcc)
    // and each touching candidate's contact index (cidx); per pass each touching transposed slot's rank in its mirror
    // row (tRank).
    // cidxGen[p]: the build (candBuilds) whose candidates candAcc[p] / cidx[p] index, 0 when the list at parity p
    // came another way; a pass maps its history through them when that is the current build's.
    DBuf<uint32_t> candA, tStart, tC, tPos, tRank, cidx[2];
    DBuf<uint8_t> candAcc[2];
    uint64_t cidxGen[2] = {0, 0};
    double skin() const {
        double s;
        if (skinSet >= 0.0) s = skinSet;
        else if (cellPassesLeft > 0) s = 0.0;
        else s = skinAuto > 0.0 ? skinAuto : 0.2 * (double)meanRadius;
        // a skin too thin for the build positions' fixed-point unit (a world very large for its spheres) would find
        // every pass stale: the cell sweep runs instead
        return 0.5 * s > SKIN_SLACK + 4.0 * posUnit ? s : 0.0;
    }
    int pairMode = -1;  // pair sweep mode (launch_cell_pairs); -1: chosen from the radius spread
    uint32_t pairBlock = 256;  // pair sweep workgroup size
    double cycleHeadroom = 1.25;  // capacity of a captured interval: headroom x the current counts + slack
    uint64_t cycleSlack = 1024;
    // force kernel specialised at run time (hipRTC); inactive: the ahead-of-time kernel of the built-in model
    JitForce jit;
    bool jitActive = false, jitStale = false, jitCustom = false;
    // owner update specialised with motion strategies / family change rules (inactive: the ahead-of-time kernel)
    JitOwner jitOwner;
    bool ownerJitActive = false;
    std::vector<std::pair<std::string, std::string>> ownerValues;
    std::string jitModel;
    std::vector<std::string> jitWildcards;
    std::vector<std::pair<std::string, std::string>> jitOverrides;
    // inspectors: one hipRTC module per region code; block partials (+ the result) and the clump volumes
    std::unordered_map<std::string, JitInspect> inspectJit;
    DBuf<double> inspPart;
    DBuf<float> ownerVol;
    bool haveVolumes = false;
    uint32_t nS = 0, nO = 0, nClumps = 0, nA = 0, nM = 0, nAnalOwners = 0;
    uint32_t maxMatSph = 0, maxMatAnal = 0, maxMat = 0;
    uint64_t nBins = 0;
    std::vector<float> hostMass;
    std::vector<float> hostRadius;  // sphere radii (the host's copy, for the cell grid after size changes)
    bool needDetect = false;        // clumps were appended: the next step starts with a detection pass
    // scenes of a batch: per owner scene index; analytical components grouped by scene (planeStart)
    std::vector<uint32_t> hostScene, hostAnalOwner, hostAnalScene;  // analScene empty: the owner's scene
    uint32_t nScenes = 1;
    bool scenesDirty = true;
    DBuf<uint32_t> planeStart;

    // static data
    DBuf<SphereRec> sph;
    DBuf<OwnerRec> own;
    DBuf<float4> moi;
    DBuf<uint32_t> ownerSphStart;
    DBuf<deme_anal_obj> anal;
    DBuf<PlanePose> pose;
    DBuf<deme_material> mats;
    DBuf<deme_family_rule> rules;
    DBuf<uint8_t> mask;
    DBuf<float> acc;  // 6 x nO

    // contact detection
    DBuf<uint32_t> binCount, binStart;
    DBuf<uint16_t> numBins;
    DBuf<uint32_t> numAnal;
    DBuf<CdRec> rec;
    DBuf<uint2> binEnt;
    DBuf<uint2> pairs;
    DBuf<uint32_t> rowCount;
    // mirror rows: per sphere, the A ids of the contacts in which it is body B (ascending), and each sphere-sphere
    // contact's slot there
    DBuf<uint32_t> bCnt, bRowStart, bPart, mirror;
    DBuf<uint64_t> scanScratch;
    DBuf<float4> sortedF;       // float copies of the cell-sorted records (the pair sweep's pre-reject)
    DBuf<uint2> scanTmp;        // gathered (row, mirror-row) counts of the fused scan
    DBuf<uint64_t> scan2Sums;   // its tile sums
    DBuf<DevCounters> ctr;
    DevCounters* hctr = nullptr;  // pinned mirror
    hipEvent_t evCounts = nullptr;  // recorded after a detection pass's counters are copied to hctr

    // double-buffered contact arrays (index cur = current list, cur ^ 1 = previous list)
    DBuf<uint32_t> rowStart[2], idB[2];
    DBuf<uint8_t> ctype[2];
    DBuf<float4> hist[2];
    DBuf<uint32_t> idA, map;
    DBuf<ForceRec> frc;           // full contact records, kept by deme_force only
    DBuf<float> aTerm, bTerm, gTerm;  // owner terms: A side per contact, B side per mirror slot / per plane contact
    DBuf<uint8_t> aOn, bOn;           // term written (the pair touches) per contact / per mirror slot
    bool forcesFull = false;
    // adjoint: checkpoints at window starts (+ the final state) and the recorded steps of one window
    struct AdjCk {
        DBuf<OwnerRec> own;
        DBuf<uint32_t> rowStart, idB;
        DBuf<uint8_t> ctype;
        DBuf<float4> hist;
        DBuf<uint32_t> idA, mirror, bRowStart;  // the list's derived rows, kept for the final checkpoint only
        DBuf<float> acc;                          // accelerations of the last step, final checkpoint only
        uint64_t C = 0, nPairs = 0;
        bool haveList = false, full = false;
        double time = 0.0;
        uint64_t nSteps = 0;
    };
    std::vector<AdjCk> adjCk;
    std::vector<DBuf<OwnerRec>> adjOwn;
    std::vector<DBuf<float4>> adjHist;
    std::vector<DBuf<float>> adjAcc;
    int adjWindow = 0, adjStepsTotal = 0;
    bool adjReady = false, singleSpheres = false;
    std::vector<deme_material> hostMats;
    DBuf<float> lam, lAcc, adjA, adjB, matPart, matGrad;
    DBuf<float4> lDt[2], lDtPrev;
    bool termsReady = false;  // owner terms belong to the current contact list
    DBuf<float> analPartials, analAcc;
    DBuf<double> clock;  // simulation time read by the owner update's motion rules (written before each such update)
    // cell-list detection
    CellGrid grid{};
    uint64_t nCells = 0;
    float maxRadius = 0.f, meanRadius = 0.f;
    uint32_t slabCap = 32;
    double cellScale = 1.0;  // cell edge = cellScale x the largest interaction distance (>= 1)
    DBuf<uint32_t> cellCount, cellStart, cellOf, slotOf, posOf, aCnt, survSlab, rowSlab, mirSlab;
    DBuf<CdRec> sorted;
    DBuf<uint2> sortedMeta;
    DBuf<unsigned long long> planeBits;
    int cur = 0;
    bool havePrev = false, contactsReady = false, fresh = false;
    uint64_t P = 0, C = 0, Cprev = 0, nPairs = 0, nActive = 0, maxOcc = 0, nGeo = 0;
    int cdStage = 0;  // 0 none, 1 counted, 2 filled, 3 pairs, 4 done

    double time = 0.0;
    uint64_t nSteps = 0;

    // per-kernel timing
    bool timing = false;
    std::vector<std::string> knames;
    std::map<std::string, int> kidx;
    std::vector<double> ktotal;
    std::vector<uint64_t> kcount;
    std::vector<hipEvent_t> evPool;
    std::vector<TimedLaunch> pending;

    int fail(int code, const std::string& msg) {
        err = msg;
        return code;
    }
    int fail_hip(hipError_t e, const char* where) {
        err = std::string("HIP error in ") + where + ": " + hipGetErrorString(e);
        return DEME_ERR_HIP;
    }
};

namespace {

hipEvent_t take_event(deme_ctx* c) {
    if (c->evPool.empty()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    hipEvent_t e = c->evPool.back();
    c->evPool.pop_back();
    return e;
}

int kernel_id(deme_ctx* c, const char* name) {
    auto it = c->kidx.find(name);
    if (it != c->kidx.end()) return it->second;
    int id = (int)c->knames.size();
    c->knames.push_back(name);
    c->kidx[name] = id;
    c->ktotal.push_back(0.0);
    c->kcount.push_back(0);
    return id;
}

hipError_t drain_timing(deme_ctx* c) {
    if (c->pending.empty()) return hipSuccess;
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return e;
    for (auto& t : c->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.s, t.e) == hipSuccess) {
            c->ktotal[t.kid] += ms;
            c->kcount[t.kid]++;
        }
        c->evPool.push_back(t.s);
        c->evPool.push_back(t.e);
    }
    c->pending.clear();
    return hipSuccess;
}

struct KTimer {
    deme_ctx* c;
    int kid = -1;
    hipEvent_t s = nullptr, e = nullptr;
    KTimer(deme_ctx* ctx, const char* name) : c(ctx) {
        if (!c->timing) return;
        kid = kernel_id(c, name);
        s = take_event(c);
        e = take_event(c);
        if (s && e) (void)hipEventRecord(s, c->stream);
    }
    ~KTimer() {
        if (kid < 0 || !s || !e) return;
        (void)hipEventRecord(e, c->stream);
        c->pending.push_back({kid, s, e});
        if (c->pending.size() > 8192) (void)drain_timing(c);
    }
};

#define DEME_LAUNCH(ctx, name, expr)                          \
    do {                                                       \
        KTimer kt_(ctx, name);                                 \
        hipError_t e_ = (expr);                                \
        if (e_ != hipSuccess) return (ctx)->fail_hip(e_, name); \
    } while (0)

#define DEME_HIP(ctx, expr)                                       \
    do {                                                           \
        hipError_t e_ = (expr);                                    \
        if (e_ != hipSuccess) return (ctx)->fail_hip(e_, #expr);   \
    } while (0)

// Analytical components of scene k are planeStart[k] .. planeStart[k + 1] - 1: components must come grouped by the
// scene of their owner, at most 64 per scene (the per-sphere plane bitset).
int build_scene_planes(deme_ctx* c) {
    if (!c->scenesDirty) return DEME_OK;
    const uint32_t nSc = c->nScenes;
    std::vector<uint32_t> start(nSc + 1, 0);
    uint32_t prev = 0;
    for (uint32_t a = 0; a < c->nA; a++) {
        const uint32_t o = c->hostAnalOwner[a];
        const uint32_t sc = a < c->hostAnalScene.size() ? c->hostAnalScene[a] : (o < c->hostScene.size() ? c->hostScene[o] : 0);
        if (sc >= nSc) return c->fail(DEME_ERR_INVALID_ARGUMENT, "an analytical component's owner has no scene");
        if (sc < prev) return c->fail(DEME_ERR_INVALID_ARGUMENT, "analytical components must be grouped by scene");
        prev = sc;
        start[sc + 1]++;
    }
    for (uint32_t k = 0; k < nSc; k++) {
        if (start[k + 1] > 64) return c->fail(DEME_ERR_UNSUPPORTED, "more than 64 analytical components in one scene");
        start[k + 1] += start[k];
    }
    DEME_HIP(c, c->planeStart.ensure(nSc + 1, 1.0));
    DEME_HIP(c, hipMemcpyAsync(c->planeStart.p, start.data(), (nSc + 1) * sizeof(uint32_t), hipMemcpyHostToDevice,
                               c->stream));
    DEME_HIP(c, hipStreamSynchronize(c->stream));
    c->kp.planeStart = c->planeStart.p;
    c->kp.nScenes = nSc;
    c->scenesDirty = false;
    return DEME_OK;
}

int check_ready(deme_ctx* c) {
    if (!c->haveParams) return c->fail(DEME_ERR_NOT_READY, "simulation parameters not set (deme_set_params)");
    if (!c->haveTemplates) return c->fail(DEME_ERR_NOT_READY, "templates not uploaded (deme_upload_templates)");
    if (!c->haveState) return c->fail(DEME_ERR_NOT_READY, "owner state not uploaded (deme_upload_state)");
    if (!c->haveMaterials) return c->fail(DEME_ERR_NOT_READY, "materials not uploaded (deme_upload_materials)");
    if (!c->haveFamilies) return c->fail(DEME_ERR_NOT_READY, "family data not uploaded (deme_upload_families)");
    if (c->maxMat >= c->nM) return c->fail(DEME_ERR_INVALID_ARGUMENT, "a material index exceeds the material table");
    return build_scene_planes(c);
}

int readback_counters(deme_ctx* c) {
    DEME_HIP(c, hipMemcpyAsync(c->hctr, c->ctr.p, DEV_COUNTERS_HOST_BYTES, hipMemcpyDeviceToHost, c->stream));
    DEME_HIP(c, hipStreamSynchronize(c->stream));
    return DEME_OK;
}

}  // namespace

extern "C" {

int deme_device_count(int* n) {
    int k = 0;
    hipError_t e = hipGetDeviceCount(&k);
    if (e != hipSuccess) k = 0;
    *n = k;
    return k > 0 ? DEME_OK : DEME_ERR_NO_DEVICE;
}

int deme_ctx_create(int device, void* stream, deme_ctx** out) {
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return DEME_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return DEME_ERR_INVALID_ARGUMENT;
    deme_ctx* c = new (std::nothrow) deme_ctx();
    if (!c) return DEME_ERR_INVALID_ARGUMENT;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess ||
        noalloc_scan_scratch(device, &c->noallocScratch, &c->noallocElems) != hipSuccess) {
        delete c;
        return DEME_ERR_HIP;
    }
    if (stream) {
        c->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            return DEME_ERR_HIP;
        }
        c->ownStream = true;
    }
    if (c->ctr.ensure(1, 1.0) != hipSuccess || c->clock.ensure(1, 1.0) != hipSuccess ||
        hipHostMalloc((void**)&c->hctr, sizeof(DevCounters), hipHostMallocDefault) != hipSuccess) {
        delete c;
        return DEME_ERR_HIP;
    }
    std::memset(c->hctr, 0, sizeof(DevCounters));
    if (hipEventCreateWithFlags(&c->evCounts, hipEventDisableTiming) != hipSuccess) {
        c->evCounts = nullptr;
        deme_ctx_destroy(c);
        return DEME_ERR_HIP;
    }
    *out = c;
    return DEME_OK;
}

int deme_ctx_destroy(deme_ctx* c) {
    if (!c) return DEME_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& t : c->pending) {
        (void)hipEventDestroy(t.s);
        (void)hipEventDestroy(t.e);
    }
    for (auto e : c->evPool) (void)hipEventDestroy(e);
    c->sph.release(); c->own.release(); c->moi.release(); c->ownerSphStart.release(); c->anal.release();
    c->pose.release(); c->mats.release(); c->rules.release(); c->mask.release(); c->acc.release();
    c->binCount.release(); c->binStart.release(); c->numBins.release(); c->numAnal.release(); c->rec.release();
    c->binEnt.release(); c->pairs.release(); c->rowCount.release(); c->bCnt.release(); c->bRowStart.release();
    c->bPart.release(); c->mirror.release(); c->scanScratch.release(); c->ctr.release();
    c->aTerm.release(); c->bTerm.release(); c->gTerm.release(); c->aOn.release(); c->bOn.release();
    for (int k = 0; k < 2; k++) {
        c->rowStart[k].release(); c->idB[k].release(); c->ctype[k].release(); c->hist[k].release();
    }
    c->idA.release(); c->map.release(); c->frc.release(); c->analPartials.release(); c->analAcc.release();
    c->clock.release();
    c->cellCount.release(); c->cellStart.release(); c->cellOf.release(); c->slotOf.release(); c->posOf.release();
    c->survSlab.release(); c->rowSlab.release(); c->mirSlab.release(); c->sorted.release(); c->planeBits.release();
    c->aCnt.release(); c->sortedMeta.release();
    c->pos.release(); c->pos0.release(); c->meta.release(); c->candStart.release(); c->candB.release();
    c->candA.release(); c->tStart.release(); c->tC.release(); c->tPos.release(); c->tRank.release();
    c->dispRef.release();
    for (int k = 0; k < 2; k++) {
        c->cidx[k].release();
        c->candAcc[k].release();
    }
    for (auto& k : c->adjCk) {
        k.own.release(); k.rowStart.release(); k.idB.release(); k.ctype.release(); k.hist.release();
        k.idA.release(); k.mirror.release(); k.bRowStart.release(); k.acc.release();
    }
    for (auto& b : c->adjOwn) b.release();
    for (auto& b : c->adjHist) b.release();
    for (auto& b : c->adjAcc) b.release();
    c->lam.release(); c->lAcc.release(); c->adjA.release(); c->adjB.release(); c->matPart.release();
    c->matGrad.release(); c->lDt[0].release(); c->lDt[1].release(); c->lDtPrev.release();
    c->jit.release();
    c->jitOwner.release();
    for (auto& kv : c->inspectJit) kv.second.release();
    c->inspPart.release(); c->ownerVol.release(); c->scanTmp.release(); c->scan2Sums.release(); c->sortedF.release();
    if (c->stepExec) (void)hipGraphExecDestroy(c->stepExec);
    for (auto& cy : c->cycle)
        if (cy.exec) (void)hipGraphExecDestroy(cy.exec);
    if (c->hctr) (void)hipHostFree(c->hctr);
    if (c->evCounts) (void)hipEventDestroy(c->evCounts);
    if (c->ownStream) (void)hipStreamDestroy(c->stream);
    delete c;
    return DEME_OK;
}

const char* deme_last_error(const deme_ctx* c) { return c ? c->err.c_str() : "null context"; }

int deme_get_stream(const deme_ctx* c, void** s) {
    *s = (void*)c->stream;
    return DEME_OK;
}

int deme_sync(deme_ctx* c) {
    DEME_HIP(c, hipStreamSynchronize(c->stream));
    return DEME_OK;
}

int deme_set_params(deme_ctx* c, const deme_sim_params* p) {
    c->graphGen++;
    c->candValid = false;  // positions, radii or the margin changed: candidate lists are rebuilt
    c->adjReady = false;  // recorded adjoint checkpoints no longer describe the system
    if (p->nvXp2 + p->nvYp2 + p->nvZp2 != 64) return c->fail(DEME_ERR_INVALID_ARGUMENT, "nvXp2 + nvYp2 + nvZp2 != 64");
    if (!(p->binSize > 0.0)) return c->fail(DEME_ERR_INVALID_ARGUMENT, "bin size must be positive");
    if (!(p->l > 0.0) || !(p->voxelSize > 0.0)) return c->fail(DEME_ERR_INVALID_ARGUMENT, "l / voxelSize must be > 0");
    const uint64_t nb = (uint64_t)p->nbX * p->nbY * p->nbZ;
    if (nb == 0 || nb > 0xFFFFFFFFull)
        return c->fail(DEME_ERR_CAPACITY, "number of bins exceeds binID_t (2^32 - 1); use a larger bin size");
    c->sp = *p;
    KParams& k = c->kp;
    k.l = p->l;
    k.voxelSize = p->voxelSize;
    k.binSize = p->binSize;
    k.nbX = p->nbX;
    k.nbY = p->nbY;
    k.nbZ = p->nbZ;
    k.nBins = (uint32_t)nb;
    k.nvX = p->nvXp2;
    k.nvY = p->nvYp2;
    // candidate-list build positions are kept in fixed point with a power-of-two unit such that +-2^31 units cover
    // twice the bin grid's largest extent
    {
        const double extent = std::max({(double)p->nbX, (double)p->nbY, (double)p->nbZ}) * p->binSize;
        const int e = std::ilogb(std::max(2.0 * extent, 1e-300)) + 1;  // 2^e > twice the extent
        c->posUnit = std::ldexp(1.0, e - 31);
    }
    k.beta = p->beta;
    k.h = p->h;
    k.Gx = p->G[0];
    k.Gy = p->G[1];
    k.Gz = p->G[2];
    k.integrator = p->integrator;
    k.nS = c->nS;
    k.nA = c->nA;
    // the per-bin arrays of the staged path are sized on its first use (deme_cd_bin_count)
    c->nBins = nb;
    c->haveParams = true;
    if (c->jitActive) c->jitStale = true;  // the specialised kernel holds the old constants
    return DEME_OK;
}

int deme_upload_templates(deme_ctx* c, uint32_t nS, const uint32_t* sphOwner, const float* sphRadius,
                          const float* sphRelPos, const uint8_t* sphMat, uint32_t nO, uint32_t nClumps,
                          const float* ownerMass, const float* ownerMoi) {
    c->graphGen++;
    c->candValid = false;  // positions, radii or the margin changed: candidate lists are rebuilt
    c->adjReady = false;  // recorded adjoint checkpoints no longer describe the system
    if (nClumps > nO) return c->fail(DEME_ERR_INVALID_ARGUMENT, "nClumps > nOwners");
    if (nO - nClumps > MAX_ANAL_OWNERS)
        return c->fail(DEME_ERR_UNSUPPORTED, "more than 8 external-object owners are not supported");
    std::vector<SphereRec> hs(nS);
    std::vector<uint32_t> start(nClumps + 1, 0);
    uint32_t maxMat = 0;
    float maxR = 0.f;
    for (uint32_t s = 0; s < nS; s++) {
        maxMat = std::max<uint32_t>(maxMat, sphMat[s]);
        if (sphOwner[s] >= nClumps) return c->fail(DEME_ERR_INVALID_ARGUMENT, "sphere owner is not a clump");
        if (s > 0 && sphOwner[s] < sphOwner[s - 1])
            return c->fail(DEME_ERR_INVALID_ARGUMENT, "spheres must be grouped by owner in ascending order");
        hs[s].owner = sphOwner[s];
        hs[s].r = sphRadius[s];
        if (!std::isfinite(sphRadius[s]) || sphRadius[s] < 0.f)
            return c->fail(DEME_ERR_INVALID_ARGUMENT, "sphere radii must be finite and non-negative");
        maxR = std::max(maxR, sphRadius[s]);
        hs[s].rx = sphRelPos[3 * s];
        hs[s].ry = sphRelPos[3 * s + 1];
        hs[s].rz = sphRelPos[3 * s + 2];
        hs[s].mat = sphMat[s];
        hs[s].pad0 = hs[s].pad1 = 0;
        start[sphOwner[s] + 1]++;
    }
    for (uint32_t o = 0; o < nClumps; o++) start[o + 1] += start[o];
    std::vector<float4> hm(nO);
    for (uint32_t o = 0; o < nO; o++) hm[o] = make_float4(ownerMoi[3 * o], ownerMoi[3 * o + 1], ownerMoi[3 * o + 2], 0.f);
    c->hostMass.assign(ownerMass, ownerMass + nO);
    // the adjoint handles single-sphere owners (each sphere its own owner, centred)
    c->singleSpheres = nS == nClumps;
    for (uint32_t s = 0; s < nS && c->singleSpheres; s++)
        c->singleSpheres = sphOwner[s] == s && sphRelPos[3 * s] == 0.f && sphRelPos[3 * s + 1] == 0.f &&
                           sphRelPos[3 * s + 2] == 0.f;
    DEME_HIP(c, c->sph.ensure(nS, 1.0));
    DEME_HIP(c, c->moi.ensure(nO, 1.0));
    DEME_HIP(c, c->own.ensure(nO, 1.0));
    DEME_HIP(c, c->acc.ensure((size_t)6 * nO, 1.0));
    DEME_HIP(c, c->ownerSphStart.ensure(nClumps + 1, 1.0));
    DEME_HIP(c, c->numBins.ensure(nS, 1.0));
    DEME_HIP(c, c->numAnal.ensure(nS, 1.0));
    DEME_HIP(c, c->rec.ensure(nS, 1.0));
    DEME_HIP(c, c->rowCount.ensure(nS, 1.0));
    DEME_HIP(c, c->bCnt.ensure(nS + 1, 1.0));
    DEME_HIP(c, c->bRowStart.ensure(nS + 1, 1.0));
    for (int k = 0; k < 2; k++) DEME_HIP(c, c->rowStart[k].ensure(nS + 1, 1.0));
    DEME_HIP(c, c->scanScratch.ensure(scan_scratch_elems(std::max<uint64_t>((uint64_t)nS, 1u << 22)), 1.0));
    if (nS) DEME_HIP(c, hipMemcpyAsync(c->sph.p, hs.data(), nS * sizeof(SphereRec), hipMemcpyHostToDevice, c->stream));
    if (nO) DEME_HIP(c, hipMemcpyAsync(c->moi.p, hm.data(), nO * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    DEME_HIP(c, hipMemcpyAsync(c->ownerSphStart.p, start.data(), (nClumps + 1) * sizeof(uint32_t),
                               hipMemcpyHostToDevice, c->stream));
    DEME_HIP(c, hipMemsetAsync(c->rowCount.p, 0, std::max<size_t>(nS, 1) * sizeof(uint32_t), c->stream));
    DEME_HIP(c, hipMemsetAsync(c->bCnt.p, 0, (nS + 1) * sizeof(uint32_t), c->stream));
    DEME_HIP(c, hipMemsetAsync(c->acc.p, 0, std::max<size_t>((size_t)6 * nO, 1) * sizeof(float), c->stream));
    DEME_HIP(c, hipStreamSynchronize(c->stream));
    c->nS = nS;
    c->kp.nS = nS;
    c->haveVolumes = false;  // per clump: re-uploaded with the templates
    c->nO = nO;
    c->nClumps = nClumps;
    c->nAnalOwners = nO - nClumps;
    c->hostScene.assign(nO, 0u);
    c->nScenes = 1;
    c->scenesDirty = true;
    c->maxMatSph = maxMat;
    c->maxMat = std::max(c->maxMatSph, c->maxMatAnal);
    c->maxRadius = maxR;
    c->hostRadius.assign(sphRadius, sphRadius + nS);
    c->meanRadius = mean_radius(c->hostRadius);
    c->needDetect = false;
    c->haveTemplates = true;
    c->haveState = false;
    c->havePrev = c->contactsReady = c->fresh = false;
    c->C = c->Cprev = 0;
    return DEME_OK;
}

int deme_upload_materials(deme_ctx* c, uint32_t n, const deme_material* m) {
    c->graphGen++;
    c->adjReady = false;  // recorded adjoint checkpoints no longer describe the system
    if (n == 0 || n > 256) return c->fail(DEME_ERR_INVALID_ARGUMENT, "1..256 materials required");
    DEME_HIP(c, c->mats.ensure(n, 1.0));
    DEME_HIP(c, hipMemcpyAsync(c->mats.p, m, n * sizeof(deme_material), hipMemcpyHostToDevice, c->stream));
    DEME_HIP(c, hipStreamSynchronize(c->stream));
    c->nM = n;
    c->hostMats.assign(m, m + n);
    c->haveMaterials = true;
    return DEME_OK;
}

int deme_upload_analytical(deme_ctx* c, uint32_t n, const deme_anal_obj* objs) {
    c->graphGen++;
    c->adjReady = false;  // recorded adjoint checkpoints no longer describe the system
    if (!c->haveTemplates) return c->fail(DEME_ERR_NOT_READY, "upload templates before analytical objects");
    if (n > 65535) return c->fail(DEME_ERR_UNSUPPORTED, "more than 65535 analytical components");
    uint32_t maxMat = 0;
    for (uint32_t a = 0; a < n; a++) {
        if (objs[a].owner < c->nClumps || objs[a].owner >= c->nO)
            return c->fail(DEME_ERR_INVALID_ARGUMENT, "analytical component owner must be an external-object owner");
        maxMat = std::max<uint32_t>(maxMat, objs[a].material);
    }
    c->maxMatAnal = maxMat;
    c->maxMat = std::max(c->maxMatSph, c->maxMatAnal);
    DEME_HIP(c, c->anal.ensure(std::max<uint32_t>(n, 1), 1.0));
    DEME_HIP(c, c->pose.ensure(std::max<uint32_t>(n, 1), 1.0));
    if (n) DEME_HIP(c, hipMemcpyAsync(c->anal.p, objs, n * sizeof(deme_anal_obj), hipMemcpyHostToDevice, c->stream));
    DEME_HIP(c, hipStreamSynchronize(c->stream));
    c->nA = n;
    c->kp.nA = n;
    c->hostAnalOwner.resize(n);
    for (uint32_t a = 0; a < n; a++) c->hostAnalOwner[a] = objs[a].owner;
    c->hostAnalScene.clear();  // components changed: their scenes follow their owners until set again
    c->scenesDirty = true;
    return DEME_OK;
}

int deme_upload_scenes(deme_ctx* c, uint32_t nScenes, const uint32_t* ownerScene, const uint32_t* analScene) {
    c->graphGen++;
    c->candValid = false;  // positions, radii or the margin changed: candidate lists are rebuilt
    if (!c->haveTemplates) return c->fail(DEME_ERR_NOT_READY, "upload templates before the scenes");
    if (nScenes == 0 || !ownerScene) return c->fail(DEME_ERR_INVALID_ARGUMENT, "need at least one scene");
    for (uint32_t o = 0; o < c->nO; o++)
        if (ownerScene[o] >= nScenes) return c->fail(DEME_ERR_INVALID_ARGUMENT, "owner scene index out of range");
    if (analScene)
        for (uint32_t a = 0; a < c->nA; a++)
            if (analScene[a] >= nScenes) return c->fail(DEME_ERR_INVALID_ARGUMENT, "component scene index out of range");
    c->hostScene.assign(ownerScene, ownerScene + c->nO);
    if (analScene) c->hostAnalScene.assign(analScene, analScene + c->nA);
    else c->hostAnalScene.clear();
    c->nScenes = nScenes;
    c->scenesDirty = true;
    if (c->haveState && c->nO) {
        // carry the scene into the resident owner records
        std::vector<OwnerRec> h(c->nO);
        DEME_HIP(c, hipMemcpyAsync(h.data(), c->own.p, c->nO * sizeof(OwnerRec), hipMemcpyDeviceToHost, c->stream));
        DEME_HIP(c, hipStreamSynchronize(c->stream));
        for (uint32_t o = 0; o < c->nO; o++) h[o].scene = ownerScene[o];
        DEME_HIP(c, hipMemcpyAsync(c->own.p, h.data(), c->nO * sizeof(OwnerRec), hipMemcpyHostToDevice, c->stream));
        DEME_HIP(c, hipStreamSynchronize(c->stream));
    }
    return DEME_OK;
}

int deme_upload_families(deme_ctx* c, const uint8_t* mask, const deme_family_rule* rules) {
    c->graphGen++;
    c->adjReady = false;  // recorded adjoint checkpoints no longer describe the system
    DEME_HIP(c, c->mask.ensure(DEME_NUM_MASK_ENTRIES, 1.0));
    DEME_HIP(c, c->rules.ensure(DEME_NUM_FAMILIES, 1.0));
    if (!rules && !c->haveFamilies) return c->fail(DEME_ERR_INVALID_ARGUMENT, "family rules required on first upload");
    DEME_HIP(c, hipMemcpyAsync(c->mask.p, mask, DEME_NUM_MASK_ENTRIES, hipMemcpyHostToDevice, c->stream));
    if (rules)
        DEME_HIP(c, hipMemcpyAsync(c->rules.p, rules, DEME_NUM_FAMILIES * sizeof(deme_family_rule),
                                   hipMemcpyHostToDevice, c->stream));
    DEME_HIP(c, hipStreamSynchronize(c->stream));
    c->maskAny = false;  // an all-allow mask lets the pair pass skip the family lookups
    for (uint32_t k = 0; k < DEME_NUM_MASK_ENTRIES; k++) c->maskAny |= mask[k] != DEME_DONT_PREVENT_CONTACT;
    c->haveFamilies = true;
    return DEME_OK;
}

int deme_upload_state(deme_ctx* c, const deme_owner_state* s) {
    c->candValid = false;  // positions, radii or the margin changed: candidate lists are rebuilt
    c->adjReady = false;  // recorded adjoint checkpoints no longer describe the system
    if (!c->haveTemplates) return c->fail(DEME_ERR_NOT_READY, "upload templates before the owner state");
    const uint32_t n = c->nO;
    std::vector<OwnerRec> h(n);
    if (c->haveState && n) {
        DEME_HIP(c, hipMemcpyAsync(h.data(), c->own.p, n * sizeof(OwnerRec), hipMemcpyDeviceToHost, c->stream));
        DEME_HIP(c, hipStreamSynchronize(c->stream));
    } else {
        std::memset(h.data(), 0, n * sizeof(OwnerRec));
        for (uint32_t o = 0; o < n; o++) h[o].qw = 1.f;
    }
    for (uint32_t o = 0; o < n; o++) {
        OwnerRec& r = h[o];
        if (s->voxelID) r.voxel = s->voxelID[o];
        if (s->locX) r.lx = s->locX[o];
        if (s->locY) r.ly = s->locY[o];
        if (s->locZ) r.lz = s->locZ[o];
        if (s->family) r.fam = s->family[o];
        if (s->oriQw) r.qw = s->oriQw[o];
        if (s->oriQx) r.qx = s->oriQx[o];
        if (s->oriQy) r.qy = s->oriQy[o];
        if (s->oriQz) r.qz = s->oriQz[o];
        if (s->vX) r.vx = s->vX[o];
        if (s->vY) r.vy = s->vY[o];
        if (s->vZ) r.vz = s->vZ[o];
        if (s->omgX) r.wx = s->omgX[o];
        if (s->omgY) r.wy = s->omgY[o];
        if (s->omgZ) r.wz = s->omgZ[o];
        r.mass = c->hostMass[o];
        r.pad0 = 0;
        r.scene = c->hostScene[o];
    }
    if (n) DEME_HIP(c, hipMemcpyAsync(c->own.p, h.data(), n * sizeof(OwnerRec), hipMemcpyHostToDevice, c->stream));
    // accelerations, when given
    float* accs[6] = {s->aX, s->aY, s->aZ, s->alphaX, s->alphaY, s->alphaZ};
    for (int k = 0; k < 6; k++)
        if (accs[k] && n)
            DEME_HIP(c, hipMemcpyAsync(c->acc.p + (size_t)k * n, accs[k], n * sizeof(float), hipMemcpyHostToDevice,
                                       c->stream));
    DEME_HIP(c, hipStreamSynchronize(c->stream));
    c->haveState = true;
    return DEME_OK;
}

int deme_clear_contacts(deme_ctx* c) {
    c->graphGen++;
    c->candValid = false;  // positions, radii or the margin changed: candidate lists are rebuilt
    c->adjReady = false;
    c->havePrev = c->contactsReady = c->fresh = false;
    c->termsReady = c->forcesFull = false;
    c->C = c->Cprev = c->nPairs = 0;
    c->nSteps = 0;
    c->time = 0.0;
    return DEME_OK;
}

int deme_download_state(deme_ctx* c, deme_owner_state* s) {
    if (!c->haveState) return c->fail(DEME_ERR_NOT_READY, "no owner state");
    const uint32_t n = c->nO;
    std::vector<OwnerRec> h(n);
    std::vector<float> a((size_t)6 * n);
    if (n) {
        DEME_HIP(c, hipMemcpyAsync(h.data(), c->own.p, n * sizeof(OwnerRec), hipMemcpyDeviceToHost, c->stream));
        DEME_HIP(c, hipMemcpyAsync(a.data(), c->acc.p, (size_t)6 * n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    }
    DEME_HIP(c, hipStreamSynchronize(c->stream));
    for (uint32_t o = 0; o < n; o++) {
        const OwnerRec& r = h[o];
        if (s->voxelID) s->voxelID[o] = r.voxel;
        if (s->locX) s->locX[o] = r.lx;
        if (s->locY) s->locY[o] = r.ly;
        if (s->locZ) s->locZ[o] = r.lz;
        if (s->family) s->family[o] = r.fam;
        if (s->oriQw) s->oriQw[o] = r.qw;
        if (s->oriQx) s->oriQx[o] = r.qx;
        if (s->oriQy) s->oriQy[o] = r.qy;
        if (s->oriQz) s->oriQz[o] = r.qz;
        if (s->vX) s->vX[o] = r.vx;
        if (s->vY) s->vY[o] = r.vy;
        if (s->vZ) s->vZ[o] = r.vz;
        if (s->omgX) s->omgX[o] = r.wx;
        if (s->omgY) s->omgY[o] = r.wy;
        if (s->omgZ) s->omgZ[o] = r.wz;
    }
    float* accs[6] = {s->aX, s->aY, s->aZ, s->alphaX, s->alphaY, s->alphaZ};
    for (int k = 0; k < 6; k++)
        if (accs[k]) std::memcpy(accs[k], a.data() + (size_t)k * n, n * sizeof(float));
    return DEME_OK;
}

// ---------------------------------------------------------------------------------------------------------------
// Contact detection stages
// ---------------------------------------------------------------------------------------------------------------
int deme_cd_bin_count(deme_ctx* c) {
    int r = check_ready(c);
    if (r) return r;
    DEME_HIP(c, c->binCount.ensure(c->nBins, 1.0));
    DEME_HIP(c, c->binStart.ensure(c->nBins + 1, 1.0));
    DEME_HIP(c, c->scanScratch.ensure(scan_scratch_elems(std::max<uint64_t>({c->nBins, (uint64_t)c->nS, 1u << 22})), 1.0));
    DEME_HIP(c, hipMemsetAsync(c->ctr.p, 0, sizeof(DevCounters), c->stream));
    DEME_LAUNCH(c, "plane_pose", launch_plane_pose(c->kp, c->anal.p, c->nA, c->own.p, c->pose.p, c->stream));
    // the occupancy histogram must start at zero whatever ran before (an aborted pass leaves counts behind)
    DEME_LAUNCH(c, "fill", launch_fill_u32(c->binCount.p, c->nBins, 0, c->stream));
    DEME_LAUNCH(c, "sphere_bins",
                launch_sphere_bins(c->kp, c->sph.p, c->own.p, c->nS, c->pose.p, c->nA, c->mask.p, c->binCount.p,
                                   c->numBins.p, c->numAnal.p, c->rec.p, c->ctr.p, c->stream));
    DEME_LAUNCH(c, "bin_stats", launch_bin_stats(c->binCount.p, c->nBins, c->ctr.p, c->stream));
    c->cdStage = 1;
    return DEME_OK;
}

int deme_cd_bin_fill(deme_ctx* c) {
    if (c->cdStage != 1) return c->fail(DEME_ERR_NOT_READY, "deme_cd_bin_count must run first");
    DEME_LAUNCH(c, "scan_bins",
                scan_u32(c->binCount.p, nullptr, c->nBins, c->binStart.p, &c->ctr.p->total[0], c->scanScratch.p,
                         c->stream));
    int r = readback_counters(c);
    if (r) return r;
    const DevCounters& h = *c->hctr;
    if (h.err || h.maxOcc > DEME_MAX_SPHERES_PER_BIN) {
        // discard the histogram so the next pass starts from zero
        DEME_LAUNCH(c, "fill", launch_fill_u32(c->binCount.p, c->nBins, 0, c->stream));
        DEME_HIP(c, hipStreamSynchronize(c->stream));
        c->cdStage = 0;
        if (h.err & ERRBIT_CAPACITY)
            return c->fail(DEME_ERR_CAPACITY, "a sphere touches more than 65535 bins (binsSphereTouches_t)");
        if (h.err & ERRBIT_OUTSIDE) return c->fail(DEME_ERR_CAPACITY, "a sphere left the bin grid of the simulation world");
        char buf[160];
        std::snprintf(buf, sizeof(buf), "a bin contains %u sphere components, exceeding maximum allowance (%u)",
                      h.maxOcc, DEME_MAX_SPHERES_PER_BIN);
        return c->fail(DEME_ERR_BIN_OVERFLOW, buf);
    }
    c->P = h.total[0];
    c->nActive = h.nActive;
    c->maxOcc = h.maxOcc;
    if (c->P > 0xFFFFFFFFull) {
        DEME_LAUNCH(c, "fill", launch_fill_u32(c->binCount.p, c->nBins, 0, c->stream));
        c->cdStage = 0;
        return c->fail(DEME_ERR_CAPACITY, "bin-sphere touch count exceeds 2^32 - 1");
    }
    DEME_HIP(c, c->binEnt.ensure(c->P));
    DEME_LAUNCH(c, "bin_scatter",
                launch_bin_scatter(c->kp, c->rec.p, c->sph.p, c->nS, c->binStart.p, c->binCount.p, c->binEnt.p,
                                   c->ctr.p, c->stream));
    c->cdStage = 2;
    return DEME_OK;
}

int deme_cd_pairs(deme_ctx* c) {
    if (c->cdStage != 2) return c->fail(DEME_ERR_NOT_READY, "deme_cd_bin_fill must run first");
    uint64_t cap = std::max<uint64_t>({(uint64_t)(c->nPairs * 1.25) + 4096, c->pairs.cap, c->P / 2 + 4096});
    // This stage accumulates per-sphere row and mirror-row counts with atomics; other passes leave their own values
    // in both arrays, so they are cleared here rather than assumed to be zero.
    DEME_LAUNCH(c, "fill", launch_fill_u32(c->rowCount.p, c->nS, 0, c->stream));
    DEME_LAUNCH(c, "fill", launch_fill_u32(c->bCnt.p, c->nS, 0, c->stream));
    for (int attempt = 0; attempt < 3; attempt++) {
        DEME_HIP(c, c->pairs.ensure(cap, 1.0));
        DEME_HIP(c, hipMemsetAsync(&c->ctr.p->pairCount, 0, sizeof(uint32_t), c->stream));
        DEME_LAUNCH(c, "bin_pairs",
                    launch_bin_pairs(c->kp, c->binEnt.p, c->P, c->binStart.p, c->rec.p, c->sph.p, c->own.p, c->mask.p,
                                     c->pairs.p, (uint32_t)std::min<uint64_t>(c->pairs.cap, 0xFFFFFFFFull), c->ctr.p,
                                     c->rowCount.p, c->bCnt.p, c->stream));
        int r = readback_counters(c);
        if (r) return r;
        const uint64_t found = c->hctr->pairCount;
        if (found <= c->pairs.cap) {
            c->nPairs = found;
            c->cdStage = 3;
            return DEME_OK;
        }
        // capacity exceeded: clear the per-row counts and redo the pass with room for every hit
        DEME_LAUNCH(c, "fill", launch_fill_u32(c->rowCount.p, c->nS, 0, c->stream));
        DEME_LAUNCH(c, "fill", launch_fill_u32(c->bCnt.p, c->nS, 0, c->stream));
        cap = found + found / 4 + 4096;
    }
    c->cdStage = 0;
    return c->fail(DEME_ERR_CAPACITY, "pair buffer could not be sized");
}

int deme_cd_sort_remap(deme_ctx* c) {
    if (c->cdStage != 3) return c->fail(DEME_ERR_NOT_READY, "deme_cd_pairs must run first");
    // the current list becomes the previous one
    if (c->contactsReady) {
        c->cur ^= 1;
        c->havePrev = true;
        c->Cprev = c->C;
    }
    const int cu = c->cur, pv = cu ^ 1;
    c->cidxGen[cu] = 0;
    DEME_LAUNCH(c, "scan_rows",
                scan_u32(c->rowCount.p, c->numAnal.p, c->nS, c->rowStart[cu].p, &c->ctr.p->total[1], c->scanScratch.p,
                         c->stream));
    int r = readback_counters(c);
    if (r) return r;
    const uint64_t C = c->hctr->total[1];
    if (C > 0xFFFFFFFFull) return c->fail(DEME_ERR_CAPACITY, "contact count exceeds contactPairs_t (2^32 - 1)");
    c->C = C;
    c->nGeo = C - c->nPairs;
    DEME_HIP(c, c->idB[cu].ensure(C));
    DEME_HIP(c, c->ctype[cu].ensure(C));
    DEME_HIP(c, c->hist[cu].ensure(C));
    DEME_HIP(c, c->idA.ensure(C));
    DEME_HIP(c, c->map.ensure(C));
    DEME_HIP(c, c->mirror.ensure(C));
    DEME_HIP(c, c->bPart.ensure(c->nPairs));
    DEME_HIP(c, c->aTerm.ensure((size_t)6 * C));
    DEME_HIP(c, c->aOn.ensure(C));
    DEME_HIP(c, c->gTerm.ensure((size_t)6 * C));
    DEME_HIP(c, c->bTerm.ensure((size_t)6 * c->nPairs));
    DEME_HIP(c, c->bOn.ensure(c->nPairs));
    c->termsReady = false;
    DEME_LAUNCH(c, "row_fill",
                launch_row_fill(c->pairs.p, (uint32_t)c->nPairs, c->rowStart[cu].p, c->numAnal.p, c->rowCount.p,
                                c->idB[cu].p, c->ctype[cu].p, c->stream));
    DEME_LAUNCH(c, "geo_fill",
                launch_geo_fill(c->kp, c->rec.p, c->sph.p, c->own.p, c->nS, c->pose.p, c->nA, c->mask.p, c->numAnal.p,
                                c->rowStart[cu].p, c->idB[cu].p, c->ctype[cu].p, c->stream));
    DEME_LAUNCH(c, "row_finalize",
                launch_row_finalize(c->nS, c->rowStart[cu].p, c->numAnal.p, c->idA.p, c->idB[cu].p, c->stream));
    DEME_LAUNCH(c, "scan_brows",
                scan_u32(c->bCnt.p, nullptr, c->nS, c->bRowStart.p, &c->ctr.p->total[2], c->scanScratch.p,
                         c->stream));
    DEME_LAUNCH(c, "brow_fill",
                launch_brow_fill((uint32_t)C, c->nS, c->idA.p, c->idB[cu].p, c->ctype[cu].p, c->bRowStart.p,
                                 c->bCnt.p, c->bPart.p, c->stream));
    DEME_LAUNCH(c, "brow_sort", launch_brow_sort(c->nS, c->bRowStart.p, c->bPart.p, c->stream));
    DEME_LAUNCH(c, "history_map",
                launch_history_map(c->nS, c->rowStart[cu].p, c->idB[cu].p, c->ctype[cu].p,
                                   c->havePrev ? c->rowStart[pv].p : nullptr, c->havePrev ? c->idB[pv].p : nullptr,
                                   c->havePrev ? c->ctype[pv].p : nullptr, c->map.p, c->bRowStart.p, c->bPart.p,
                                   c->mirror.p, c->stream));
    c->contactsReady = true;
    c->needDetect = false;
    c->fresh = true;
    c->forcesFull = false;
    c->cdStage = 4;
    return DEME_OK;
}

// Cell grid for candidate generation: edge = largest interaction distance 2 * (r_max + beta) (slightly padded for
// the float rounding of r + beta), covering the voxel world.
static int setup_grid(deme_ctx* c) {
    if (!std::isfinite(c->maxRadius) || !(c->maxRadius > 0.f))
        return c->fail(DEME_ERR_INVALID_ARGUMENT, "sphere radii must be positive and finite");
    // with candidate lists the cells serve their builds, whose interaction distance is widened by the skin
    const double skin = c->skin() > 0.0 ? c->skin() : 0.0;
    double edge = 2.0 * ((double)c->maxRadius + (double)c->sp.beta + 0.5 * skin) * 1.0001 * c->cellScale;
    // the grid spans the user's domain (positions beyond it fall into edge cells), at most the voxel world
    auto span = [](double voxelWorld, float box) {
        return (box > 0.f && (double)box < voxelWorld) ? (double)box : voxelWorld;
    };
    const double wx = span(c->sp.voxelSize_exact * std::ldexp(1.0, c->sp.nvXp2), c->sp.boxSize[0]);
    const double wy = span(c->sp.voxelSize_exact * std::ldexp(1.0, c->sp.nvYp2), c->sp.boxSize[1]);
    const double wz = span(c->sp.voxelSize_exact * std::ldexp(1.0, c->sp.nvZp2), c->sp.boxSize[2]);
    for (;;) {
        const double nx = std::floor(wx / edge) + 1, ny = std::floor(wy / edge) + 1, nz = std::floor(wz / edge) + 1;
        if (nx * ny * nz <= (double)(1u << 28)) {
            c->grid.inv = 1.0 / edge;
            c->grid.nx = (uint32_t)nx;
            c->grid.ny = (uint32_t)ny;
            c->grid.nz = (uint32_t)nz;
            c->grid.nCells = (uint32_t)(nx * ny * nz);
            break;
        }
        edge *= 1.25;
    }
    const uint64_t nCells = (uint64_t)c->grid.nx * c->grid.ny * c->grid.nz;
    if (nCells != c->nCells) {
        DEME_HIP(c, c->cellCount.ensure(nCells, 1.0));
        DEME_HIP(c, c->cellStart.ensure(nCells + 1, 1.0));
        DEME_HIP(c, c->scanScratch.ensure(scan_scratch_elems(std::max<uint64_t>({nCells, (uint64_t)c->nS, 1u << 22})), 1.0));
        c->nCells = nCells;
    }
    return DEME_OK;
}

// Rounding bound of the pair sweep's float pre-reject: coordinates (below the world's largest extent W) are rounded to
// within 2^-24 W, so a pair's float distance errs by less than 2 * sqrt(3) * 2^-24 W plus the rounding of the float
// arithmetic on distances far below W; 16 * 2^-24 W covers both.
static float sweep_margin(const deme_sim_params& p) {
    const double W = std::ldexp(p.voxelSize_exact, (int)std::max({p.nvXp2, p.nvYp2, p.nvZp2}));
    return (float)(W * 16.0 / 16777216.0);
}

// The canonical rows and mirror rows of the cell path, then the history map against the previous list.
// skinPass: the pass ran on candidate lists (else the cell sweep).
static int enqueue_rows(deme_ctx* c, SpecGuard guard, bool skinPass) {
    const int cu = c->cur, pv = cu ^ 1;
    const uint32_t nS = c->nS;
    if (skinPass) {
        // candidate-list pass: rows and links from the candidates' indices; the map too when the previous list
        // indexes the same candidates, else (the first pass of a build) by the merge join of the rows
        const bool mapFromCand = c->havePrev && c->cidxGen[pv] == c->candBuilds;
        DEME_LAUNCH(c, "cand_fill",
                    launch_cand_fill((uint32_t)c->nCand, c->candA.p, c->candB.p, c->candStart.p, c->candAcc[cu].p,
                                     c->rowStart[cu].p, c->numAnal.p, c->tPos.p, c->tRank.p, c->bRowStart.p,
                                     mapFromCand ? c->candAcc[pv].p : nullptr, c->cidx[pv].p, c->cidx[cu].p, c->idA.p,
                                     c->idB[cu].p, c->ctype[cu].p, c->map.p, c->mirror.p, c->stream, guard));
        DEME_LAUNCH(c, "cand_planes",
                    launch_cand_planes(c->kp, c->sph.p, c->own.p, nS, c->rowStart[cu].p, c->planeBits.p,
                                       c->havePrev ? c->rowStart[pv].p : nullptr, c->havePrev ? c->idB[pv].p : nullptr,
                                       c->havePrev ? c->ctype[pv].p : nullptr, mapFromCand, c->idA.p, c->idB[cu].p,
                                       c->ctype[cu].p, c->map.p, c->mirror.p, c->stream, guard));
        if (!mapFromCand)
            DEME_LAUNCH(c, "history_map",
                        launch_history_map(nS, c->rowStart[cu].p, c->idB[cu].p, c->ctype[cu].p,
                                           c->havePrev ? c->rowStart[pv].p : nullptr,
                                           c->havePrev ? c->idB[pv].p : nullptr,
                                           c->havePrev ? c->ctype[pv].p : nullptr, c->map.p, nullptr, nullptr,
                                           nullptr, c->stream, guard));
        return DEME_OK;
    }
    DEME_LAUNCH(c, "cell_rows",
                launch_cell_rows(c->kp, c->sph.p, c->own.p, nS, c->rowSlab.p, c->mirSlab.p, c->posOf.p,
                                 c->rowStart[cu].p, c->planeBits.p, c->bRowStart.p, c->idA.p, c->idB[cu].p,
                                 c->ctype[cu].p, c->bPart.p, c->stream, guard));
    DEME_LAUNCH(c, "history_map",
                launch_history_map(nS, c->rowStart[cu].p, c->idB[cu].p, c->ctype[cu].p,
                                   c->havePrev ? c->rowStart[pv].p : nullptr, c->havePrev ? c->idB[pv].p : nullptr,
                                   c->havePrev ? c->ctype[pv].p : nullptr, c->map.p, c->bRowStart.p, c->bPart.p,
                                   c->mirror.p, c->stream, guard));
    return DEME_OK;
}

// Cell grid and the per-sphere buffers of a detection pass; *mode: the pair sweep's mode (launch_cell_pairs).
static int cd_buffers(deme_ctx* c, int* mode) {
    int r = setup_grid(c);
    if (r) return r;
    const uint32_t nS = c->nS;
    DEME_HIP(c, c->cellOf.ensure(nS));
    DEME_HIP(c, c->slotOf.ensure(nS));
    DEME_HIP(c, c->posOf.ensure(nS));
    DEME_HIP(c, c->aCnt.ensure(nS));
    DEME_HIP(c, c->sortedMeta.ensure(nS));
    DEME_HIP(c, c->sorted.ensure(nS));
    // default: the float sweep with the exact tests in the shared phase (measured fastest on near-uniform clumps,
    // with or without a drift margin, and on wide radius spreads); a mode set on the context wins
    *mode = c->pairMode >= 0 ? c->pairMode : 2;
    if (*mode != 0) DEME_HIP(c, c->sortedF.ensure(nS));
    DEME_HIP(c, c->planeBits.ensure(nS));
    DEME_HIP(c, c->scanTmp.ensure(nS));
    DEME_HIP(c, c->scan2Sums.ensure(scan2_sums_elems(nS)));
    DEME_HIP(c, c->survSlab.ensure((size_t)nS * c->slabCap, 1.0));
    DEME_HIP(c, c->rowSlab.ensure((size_t)nS * c->slabCap, 1.0));
    DEME_HIP(c, c->mirSlab.ensure((size_t)nS * c->slabCap, 1.0));
    return DEME_OK;
}

// The detection pass up to the sorted records: plane poses, the pass's counters and cell counts cleared, cell
// assignment (bin counts and plane contacts on the way), the cell scan and the counting-sort scatter.
static int enqueue_cell_pass(deme_ctx* c, int mode) {
    const uint32_t nS = c->nS;
    DEME_LAUNCH(c, "plane_pose", launch_plane_pose(c->kp, c->anal.p, c->nA, c->own.p, c->pose.p, c->stream));
    // cell counts and the pass's counters cleared by one kernel (a small memset is two fills on this runtime)
    static_assert(sizeof(DevCounters) % 4 == 0, "counters are cleared as 32-bit words");
    DEME_LAUNCH(c, "fill", launch_fill_u32(c->cellCount.p, c->nCells, 0, c->stream,
                                           reinterpret_cast<uint32_t*>(c->ctr.p), sizeof(DevCounters) / 4));
    DEME_LAUNCH(c, "cell_assign",
                launch_cell_assign(c->kp, c->grid, c->sph.p, c->own.p, nS, c->pose.p, c->nA,
                                   c->maskAny ? c->mask.p : nullptr,
                                   c->cellCount.p, c->cellOf.p, c->slotOf.p, c->numBins.p, c->numAnal.p,
                                   c->planeBits.p, c->ctr.p, c->stream));
    DEME_LAUNCH(c, "scan_cells",
                scan_u32(c->cellCount.p, nullptr, c->nCells, c->cellStart.p, &c->ctr.p->total[3], c->scanScratch.p,
                         c->stream));
    DEME_LAUNCH(c, "cell_scatter",
                launch_cell_scatter(c->kp, c->sph.p, c->own.p, nS, c->cellOf.p, c->slotOf.p, c->cellStart.p, c->sorted.p,
                                    c->sortedMeta.p, c->posOf.p, c->aCnt.p, c->bCnt.p, c->sp.beta,
                                    mode != 0 ? c->sortedF.p : nullptr, c->stream));
    return DEME_OK;
}

// The pair sweep and the row / mirror-row scans of the list at parity cu (totals in the device counters).
static int enqueue_pair_pass(deme_ctx* c, int mode, int cu) {
    const uint32_t nS = c->nS;
    DEME_LAUNCH(c, "cell_pairs",
                launch_cell_pairs(c->kp, c->grid, nS, c->sorted.p, mode != 0 ? c->sortedF.p : nullptr,
                                  sweep_margin(c->sp), mode, c->pairBlock,
                                  c->sortedMeta.p, c->cellStart.p,
                                  c->maskAny ? c->mask.p : nullptr,
                                  c->slabCap, c->survSlab.p, c->rowSlab.p, c->mirSlab.p, c->aCnt.p, c->bCnt.p,
                                  c->ctr.p, c->stream));
    // row lengths in id order: partner counts gathered from sorted positions, plus plane contacts; mirror rows
    if ((uint64_t)nS <= SCAN2_TILE * SCAN_TILE) {
        DEME_LAUNCH(c, "scan_rows",
                    scan2_u32_gather(c->aCnt.p, c->numAnal.p, c->bCnt.p, c->posOf.p, nS, c->rowStart[cu].p,
                                     c->bRowStart.p, &c->ctr.p->total[1], &c->ctr.p->total[2], c->scanTmp.p,
                                     c->scan2Sums.p, c->stream));
    } else {
        DEME_LAUNCH(c, "scan_rows",
                    scan_u32(c->aCnt.p, c->numAnal.p, nS, c->rowStart[cu].p, &c->ctr.p->total[1],
                             c->scanScratch.p, c->stream, c->posOf.p));
        DEME_LAUNCH(c, "scan_brows",
                    scan_u32(c->bCnt.p, nullptr, nS, c->bRowStart.p, &c->ctr.p->total[2], c->scanScratch.p,
                             c->stream, c->posOf.p));
    }
    return DEME_OK;
}

// The candidate-list pass is stale once some sphere has moved half the skin (less the slack) since the build.
static double stale_limit2(const deme_ctx* c) {
    // (posUnit bounds the fixed-point rounding of a build position: sqrt(3) / 2 units)
    const double h = 0.5 * c->skin() - SKIN_SLACK - c->posUnit;
    return h > 0.0 ? h * h : 0.0;
}

// Candidate lists from the current state: the cell sweep in its build mode (pairs of different owners within the
// skin of touching), rows of higher-id partners in id order (candStart / candB) and the build positions (pos0).
static int build_candidates(deme_ctx* c) {
    if (c->skinSet < 0.0 && c->candBuilds > 0 && c->candValid) {
        // the retiring lists' lifetime and size judge their skin (lists dropped for other reasons -- a host change,
        // a return from the cell sweep -- do not)
        constexpr double BUILD_PASSES = 2.7;
        const double m = (double)c->meanRadius, cur = c->skin();
        const double life = (double)std::max<uint64_t>(c->passesSinceBuild, 1);
        const double cost = (double)std::max<uint64_t>(c->nCand, 1) * (1.0 + BUILD_PASSES / life);
        const uint64_t L = c->passesSinceBuild;
        double next = cur;
        if (L <= 1) {
            next = cur * 1.5;
        } else if (L > 400) {
            next = cur * 0.75;
        } else if (L >= 30) {
            c->skinDir = 0;
        } else {
            if (c->skinCostPrev >= 0.0 && c->skinDir != 0) {
                const double rel = cost / c->skinCostPrev;
                if (rel > 1.02) c->skinDir = -c->skinDir;
                else if (rel >= 0.98) c->skinDir = 0;
            }
            if (c->skinDir == 0 && L < 8) c->skinDir = 1;
            if (c->skinDir > 0) next = cur * 1.25;
            else if (c->skinDir < 0) next = cur * 0.8;
        }
        c->skinCostPrev = cost;
        c->skinAuto = std::min(std::max(next, 0.05 * m), 0.8 * m);
    }
    c->lastLifetime = c->candBuilds > 0 ? c->passesSinceBuild : ~0ull;
    c->builtThisPass = true;
    c->passesSinceBuild = 0;
    int mode = 0;
    int r = cd_buffers(c, &mode);
    if (r) return r;
    const uint32_t nS = c->nS;
    DEME_HIP(c, c->sortedF.ensure(nS));
    DEME_HIP(c, c->candStart.ensure((size_t)nS + 1));
    DEME_HIP(c, c->tStart.ensure((size_t)nS + 1));
    DEME_HIP(c, c->pos0.ensure(nS));
    DEME_HIP(c, c->dispRef.ensure(8, 1.0));
    for (int attempt = 0;; attempt++) {
        r = enqueue_cell_pass(c, PAIR_MODE_CAND);
        if (r) return r;
        DEME_LAUNCH(c, "cand_build",
                    launch_cell_pairs(c->kp, c->grid, nS, c->sorted.p, c->sortedF.p, sweep_margin(c->sp),
                                      PAIR_MODE_CAND, c->pairBlock, c->sortedMeta.p, c->cellStart.p, nullptr,
                                      c->slabCap, c->survSlab.p, c->rowSlab.p, c->mirSlab.p, c->aCnt.p, c->bCnt.p,
                                      c->ctr.p, c->stream, c->skin()));
        DEME_LAUNCH(c, "scan_cand",
                    scan_u32(c->aCnt.p, nullptr, nS, c->candStart.p, &c->ctr.p->total[0], c->scanScratch.p,
                             c->stream, c->posOf.p));
        r = readback_counters(c);
        if (r) return r;
        const DevCounters& h = *c->hctr;
        if (h.err & ERRBIT_CAPACITY)
            return c->fail(DEME_ERR_CAPACITY, "a sphere touches more than 65535 bins (binsSphereTouches_t)");
        if (h.err) return c->fail(DEME_ERR_CAPACITY, "a sphere left the bin grid of the simulation world");
        if (h.maxOcc <= c->slabCap) break;
        if (attempt >= 2) return c->fail(DEME_ERR_CAPACITY, "candidate rows could not be sized");
        c->slabCap = (uint32_t)std::min<uint64_t>(((uint64_t)h.maxOcc * 5 / 4 + 7) & ~7ull, 65535);
        DEME_HIP(c, c->survSlab.ensure((size_t)nS * c->slabCap, 1.0));
        DEME_HIP(c, c->rowSlab.ensure((size_t)nS * c->slabCap, 1.0));
        DEME_HIP(c, c->mirSlab.ensure((size_t)nS * c->slabCap, 1.0));
    }
    c->nCand = c->hctr->total[0];
    if (c->nCand > 0xFFFFFFFFull) return c->fail(DEME_ERR_CAPACITY, "candidate count exceeds 2^32 - 1");
    const size_t nK = std::max<size_t>(c->nCand, 1);
    DEME_HIP(c, c->candA.ensure(nK));
    DEME_HIP(c, c->candB.ensure(nK));
    DEME_HIP(c, c->tC.ensure(nK));
    DEME_HIP(c, c->tPos.ensure(nK));
    DEME_HIP(c, c->tRank.ensure(nK));
    for (int k = 0; k < 2; k++) {
        DEME_HIP(c, c->candAcc[k].ensure(nK));
        DEME_HIP(c, c->cidx[k].ensure(nK));
    }
    // the build positions (in fixed point, posUnit) are the reference: no displacement, no common translation yet
    DEME_LAUNCH(c, "fill", launch_fill_u32(reinterpret_cast<uint32_t*>(c->dispRef.p), 12, 0, c->stream));
    DEME_LAUNCH(c, "cand_rows",
                launch_cand_rows(nS, c->rowSlab.p, c->candStart.p, c->candA.p, c->candB.p, c->sorted.p, c->pos0.p,
                                 1.0 / c->posUnit, c->stream));
    // the transpose, counted from the rows (cidx[0] holds each candidate's place until the first pass writes it)
    const uint32_t nK32 = (uint32_t)c->nCand;
    DEME_LAUNCH(c, "fill", launch_fill_u32(c->bCnt.p, nS, 0, c->stream));
    DEME_LAUNCH(c, "cand_tcount", launch_cand_tcount(nK32, c->candB.p, c->bCnt.p, c->cidx[0].p, c->stream));
    DEME_LAUNCH(c, "scan_cand",
                scan_u32(c->bCnt.p, nullptr, nS, c->tStart.p, &c->ctr.p->total[2], c->scanScratch.p, c->stream));
    DEME_LAUNCH(c, "cand_tscatter",
                launch_cand_tscatter(nK32, c->candB.p, c->cidx[0].p, c->tStart.p, c->tC.p, c->stream));
    DEME_LAUNCH(c, "cand_tsort", launch_cand_tsort(nS, c->tStart.p, c->tC.p, c->tPos.p, c->stream));
    c->candValid = true;
    c->rebuildNext = false;
    c->candBuilds++;
    return DEME_OK;
}

// Buffers of a candidate-list pass (the lists themselves come from build_candidates).
static int skin_buffers(deme_ctx* c) {
    const uint32_t nS = c->nS;
    DEME_HIP(c, c->pos.ensure(nS));
    DEME_HIP(c, c->meta.ensure(nS));
    DEME_HIP(c, c->aCnt.ensure(nS));
    DEME_HIP(c, c->planeBits.ensure(nS));
    DEME_HIP(c, c->scanTmp.ensure(nS));
    DEME_HIP(c, c->scan2Sums.ensure(scan2_sums_elems(nS)));
    DEME_HIP(c, c->scanScratch.ensure(scan_scratch_elems(std::max<uint64_t>(nS, 1u << 22)), 1.0));
    return DEME_OK;
}

// One candidate-list pass up to the row scans: plane poses, counters cleared, per-sphere positions (with the
// displacement since the build), the predicate on the candidates (touching flags, row counts), the mirror-row
// counts, the row and mirror-row scans.
static int enqueue_skin_pass(deme_ctx* c, int cu) {
    const uint32_t nS = c->nS;
    static_assert(sizeof(DevCounters) % 4 == 0, "counters are cleared as 32-bit words");
    DEME_LAUNCH(c, "plane_pose",
                launch_pass_start(c->kp, c->anal.p, c->nA, c->own.p, c->pose.p, reinterpret_cast<uint32_t*>(c->ctr.p),
                                  sizeof(DevCounters) / 4, c->stream));
    // (the families for the mask only when a mask is in force; no per-sphere bin counts: this path never reads them)
    DEME_LAUNCH(c, "sphere_pos",
                launch_sphere_pos(c->kp, c->sph.p, c->own.p, nS, c->pose.p, c->nA, c->maskAny ? c->mask.p : nullptr,
                                  c->pos.p, c->pos0.p, c->posUnit, c->maskAny ? c->meta.p : nullptr, nullptr,
                                  c->numAnal.p,
                                  c->planeBits.p, c->ctr.p, c->dispRef.p + 3 * (cu ^ 1), c->stream));
    DEME_LAUNCH(c, "cand_flags",
                launch_cand_flags(c->kp, (uint32_t)c->nCand, c->pos.p, c->meta.p, c->candA.p, c->candB.p,
                                  c->maskAny ? c->mask.p : nullptr, stale_limit2(c), c->candAcc[cu].p, c->ctr.p,
                                  c->dispRef.p + 3 * cu, c->stream));
    if ((uint64_t)nS <= SCAN2_TILE * SCAN_TILE) {
        // the row counts formed inside the scan's first pass
        DEME_LAUNCH(c, "scan_rows",
                    scan2_cand_rows(c->candStart.p, c->tStart.p, c->tC.p, c->candAcc[cu].p, c->numAnal.p, nS,
                                    c->tRank.p, c->rowStart[cu].p, c->bRowStart.p, &c->ctr.p->total[1], &c->ctr.p->total[2],
                                    c->scanTmp.p, c->scan2Sums.p, c->ctr.p, c->stream));
    } else {
        DEME_LAUNCH(c, "cand_count",
                    launch_cand_count(nS, c->candStart.p, c->tStart.p, c->tC.p, c->candAcc[cu].p, c->aCnt.p,
                                      c->bCnt.p, c->tRank.p, c->ctr.p, c->stream));
        DEME_LAUNCH(c, "scan_rows",
                    scan_u32(c->aCnt.p, c->numAnal.p, nS, c->rowStart[cu].p, &c->ctr.p->total[1], c->scanScratch.p,
                             c->stream));
        DEME_LAUNCH(c, "scan_brows",
                    scan_u32(c->bCnt.p, nullptr, nS, c->bRowStart.p, &c->ctr.p->total[2], c->scanScratch.p,
                             c->stream));
    }
    return DEME_OK;
}

// Bookkeeping at the end of a detection pass (candidate-list or cell sweep; disp2: the largest squared displacement
// since the build; the pass's builds, if any, are behind it). A candidate-list pass rebuilds ahead of time when the next pass would likely find the lists
// stale; with the automatic skin, lists that keep dying within 4 passes at the widest skin give way to the cell sweep
// for a while, and cell sweeps count down to trying them again.
static void end_pass(deme_ctx* c, bool skinPass, double disp2) {
    const bool built = c->builtThisPass;
    c->builtThisPass = false;
    if (skinPass) {
        c->passesSinceBuild++;
        const double lim = 0.35 * c->skin();
        c->rebuildNext = disp2 > lim * lim;
        // lists about to be rebuilt after fewer than 4 passes, or just rebuilt (found stale) after fewer than 4
        const bool dying = (c->rebuildNext && c->passesSinceBuild < 4) || (built && c->lastLifetime < 4);
        if (dying && c->skinSet < 0.0 && c->skinAuto >= 0.8 * (double)c->meanRadius * (1.0 - 1e-9)) {
            c->cellBackoff = std::min<uint32_t>(c->cellBackoff ? 2 * c->cellBackoff : 64, 4096);
            c->cellPassesLeft = c->cellBackoff;
            c->candValid = false;
            c->graphGen++;
        } else if (!c->rebuildNext && c->passesSinceBuild >= 4) {
            c->cellBackoff = 0;
        }
    } else if (c->skinSet < 0.0 && c->cellPassesLeft > 0 && --c->cellPassesLeft == 0) {
        c->candValid = false;
        c->passesSinceBuild = 100;  // the lists start afresh: no lifetime to judge them by
        c->skinCostPrev = -1.0;
        c->graphGen++;
    }
}

int deme_cd_run(deme_ctx* c) {
    c->graphGen++;
    int r = check_ready(c);
    if (r) return r;
    const bool skin = c->skin() > 0.0;
    int mode = 0;
    const uint32_t nS = c->nS;
    if (skin) {
        DEME_HIP(c, c->numBins.ensure(nS));
        DEME_HIP(c, c->numAnal.ensure(nS));
        DEME_HIP(c, c->bCnt.ensure(nS));
        DEME_HIP(c, c->bRowStart.ensure((size_t)nS + 1));
        DEME_HIP(c, c->rowStart[c->cur ^ 1].ensure((size_t)nS + 1));
        DEME_HIP(c, c->rowStart[c->cur].ensure((size_t)nS + 1));
        if (!c->candValid || c->rebuildNext) {
            r = build_candidates(c);
            if (r) return r;
        }
        r = skin_buffers(c);
        if (r) return r;
    } else {
        r = cd_buffers(c, &mode);
        if (r) return r;
        r = enqueue_cell_pass(c, mode);
        if (r) return r;
    }
    // the current list becomes the previous one
    if (c->contactsReady) {
        c->cur ^= 1;
        c->havePrev = true;
        c->Cprev = c->C;
    }
    c->contactsReady = false;
    c->cdStage = 0;
    const int cu = c->cur;
    c->cidxGen[cu] = 0;
    SpecGuard spec;
    for (int attempt = 0;; attempt++) {
        r = skin ? enqueue_skin_pass(c, cu) : enqueue_pair_pass(c, mode, cu);
        if (r) return r;
        // The counters go to the host without a stream-wide wait: the rows and the history map are enqueued behind
        // them with the buffers at hand (skipped on the device if the pass failed or the lists do not fit), so the
        // GPU keeps working while the host reads the totals.
        DEME_HIP(c, hipMemcpyAsync(c->hctr, c->ctr.p, DEV_COUNTERS_HOST_BYTES, hipMemcpyDeviceToHost, c->stream));
        DEME_HIP(c, hipEventRecord(c->evCounts, c->stream));
        spec.ctr = nullptr;
        if (c->idB[cu].p && c->ctype[cu].p && c->idA.p && c->map.p && c->mirror.p && c->bPart.p) {
            spec.ctr = c->ctr.p;
            spec.slabCap = c->slabCap;
            spec.capC = std::min({c->idB[cu].cap, c->ctype[cu].cap, c->idA.cap, c->map.cap, c->mirror.cap});
            spec.capP = c->bPart.cap;
            r = enqueue_rows(c, spec, skin);
            if (r) return r;
        }
        DEME_HIP(c, hipEventSynchronize(c->evCounts));
        const DevCounters& h = *c->hctr;
        if (h.err == ERRBIT_STALE) {
            // a sphere moved half the skin since the build: new lists, then the pass again
            if (attempt >= 3) return c->fail(DEME_ERR_CAPACITY, "candidate lists went stale right after a build");
            r = build_candidates(c);  // (the stale lists are judged by the passes they lasted)
            if (r) return r;
            continue;
        }
        if (h.err) {
            c->havePrev = false;
            if (h.err & ERRBIT_CAPACITY)
                return c->fail(DEME_ERR_CAPACITY, "a sphere touches more than 65535 bins (binsSphereTouches_t)");
            return c->fail(DEME_ERR_CAPACITY, "a sphere left the bin grid of the simulation world");
        }
        if (h.maxOcc <= c->slabCap) break;
        if (attempt >= 3) return c->fail(DEME_ERR_CAPACITY, "partner rows could not be sized");
        // a sphere has more partners than the slab row holds: widen the rows and redo the pass
        c->slabCap = (uint32_t)std::min<uint64_t>(((uint64_t)h.maxOcc * 5 / 4 + 7) & ~7ull, 65535);
        DEME_HIP(c, c->survSlab.ensure((size_t)nS * c->slabCap, 1.0));
        DEME_HIP(c, c->rowSlab.ensure((size_t)nS * c->slabCap, 1.0));
        DEME_HIP(c, c->mirSlab.ensure((size_t)nS * c->slabCap, 1.0));
        if (!skin) {
            DEME_LAUNCH(c, "fill", launch_fill_u32(c->aCnt.p, nS, 0, c->stream));
            DEME_LAUNCH(c, "fill", launch_fill_u32(c->bCnt.p, nS, 0, c->stream));
            DEME_HIP(c, hipMemsetAsync(&c->ctr.p->maxOcc, 0, sizeof(uint32_t), c->stream));
        }
    }
    const double disp2 = max_disp2(*c->hctr);
    const uint64_t C = c->hctr->total[1], nPairs = c->hctr->total[2];
    if (C > 0xFFFFFFFFull) return c->fail(DEME_ERR_CAPACITY, "contact count exceeds contactPairs_t (2^32 - 1)");
    if (nPairs > C) return c->fail(DEME_ERR_HIP, "inconsistent contact and mirror-row counts");
    c->C = C;
    c->nPairs = nPairs;
    c->nGeo = C - nPairs;
    DEME_HIP(c, c->idB[cu].ensure(C));
    DEME_HIP(c, c->ctype[cu].ensure(C));
    DEME_HIP(c, c->hist[cu].ensure(C));
    DEME_HIP(c, c->idA.ensure(C));
    DEME_HIP(c, c->map.ensure(C));
    DEME_HIP(c, c->mirror.ensure(C));
    DEME_HIP(c, c->bPart.ensure(nPairs));
    DEME_HIP(c, c->aTerm.ensure((size_t)6 * C));
    DEME_HIP(c, c->aOn.ensure(C));
    DEME_HIP(c, c->gTerm.ensure((size_t)6 * C));
    DEME_HIP(c, c->bTerm.ensure((size_t)6 * nPairs));
    DEME_HIP(c, c->bOn.ensure(nPairs));
    c->termsReady = false;
    // the rows and the map already ran unless the lists outgrew the buffers they were enqueued with
    if (!spec.ctr || C > spec.capC || nPairs > spec.capP) {
        r = enqueue_rows(c, SpecGuard(), skin);
        if (r) return r;
    }
    c->contactsReady = true;
    c->cidxGen[cu] = skin ? c->candBuilds : 0;
    end_pass(c, skin, disp2);
    c->needDetect = false;
    c->fresh = true;
    c->forcesFull = false;
    c->cdStage = 0;
    return DEME_OK;
}

int deme_prefix_scan_u32(deme_ctx* c, const uint32_t* in, uint32_t* out, uint64_t n, uint64_t* total) {
    if ((n && !in) || !out) return c->fail(DEME_ERR_INVALID_ARGUMENT, "null device array");
    DEME_HIP(c, c->scanScratch.ensure(scan_scratch_elems(n), 1.0));
    DEME_LAUNCH(c, "scan", scan_u32(in, nullptr, n, out, &c->ctr.p->total[0], c->scanScratch.p, c->stream));
    if (total) {
        unsigned long long t = 0;
        DEME_HIP(c, hipMemcpyAsync(&t, &c->ctr.p->total[0], sizeof(t), hipMemcpyDeviceToHost, c->stream));
        DEME_HIP(c, hipStreamSynchronize(c->stream));
        *total = t;
    }
    DEME_HIP(c, hipStreamSynchronize(c->stream));
    return DEME_OK;
}

int deme_prefix_scan_u32_noalloc(deme_ctx* c, const uint32_t* in, uint32_t* out, uint64_t n, uint64_t* totalDev) {
    if ((n && !in) || !out) return c->fail(DEME_ERR_INVALID_ARGUMENT, "null device array");
    if (n > NOALLOC_SCAN_MAX)
        return c->fail(DEME_ERR_INVALID_ARGUMENT,
                       "deme_prefix_scan_u32_noalloc scans at most 2^28 elements (its fixed workspace); use "
                       "deme_prefix_scan_u32 for longer arrays");
    if (!c->noallocScratch || scan_scratch_elems(n) + 1 > c->noallocElems)
        return c->fail(DEME_ERR_HIP, "the allocation-free scan's workspace is not available on this device");
    // the last slot takes the total when the caller wants none (the kernels write it unconditionally otherwise)
    unsigned long long* tot = totalDev ? reinterpret_cast<unsigned long long*>(totalDev)
                                       : reinterpret_cast<unsigned long long*>(c->noallocScratch + c->noallocElems - 1);
    // launched without the kernel timer (its event pool may grow)
    const hipError_t e = scan_u32(in, nullptr, n, out, tot, c->noallocScratch, c->stream);
    if (e != hipSuccess) return c->fail_hip(e, "scan");
    return DEME_OK;
}

int deme_download_bin_counts(deme_ctx* c, uint16_t* nb, uint8_t* na) {
    if (c->cdStage < 1) return c->fail(DEME_ERR_NOT_READY, "no detection pass");
    if (nb && c->nS) DEME_HIP(c, hipMemcpyAsync(nb, c->numBins.p, c->nS * sizeof(uint16_t), hipMemcpyDeviceToHost, c->stream));
    std::vector<uint32_t> tmp(c->nS);
    if (na && c->nS)
        DEME_HIP(c, hipMemcpyAsync(tmp.data(), c->numAnal.p, c->nS * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    DEME_HIP(c, hipStreamSynchronize(c->stream));
    if (na)
        for (uint32_t s = 0; s < c->nS; s++) na[s] = (uint8_t)tmp[s];
    return DEME_OK;
}

int deme_download_bins(deme_ctx* c, uint32_t* binStart, uint32_t* binSphere) {
    if (c->cdStage < 2) return c->fail(DEME_ERR_NOT_READY, "bins not filled");
    if (binStart)
        DEME_HIP(c, hipMemcpyAsync(binStart, c->binStart.p, (c->nBins + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                   c->stream));
    std::vector<uint2> e(c->P);
    if (binSphere && c->P)
        DEME_HIP(c, hipMemcpyAsync(e.data(), c->binEnt.p, c->P * sizeof(uint2), hipMemcpyDeviceToHost, c->stream));
    DEME_HIP(c, hipStreamSynchronize(c->stream));
    if (binSphere)
        for (uint64_t k = 0; k < c->P; k++) binSphere[k] = e[k].x;
    return DEME_OK;
}

int deme_download_contacts(deme_ctx* c, uint32_t* idA, uint32_t* idB, uint8_t* type, uint32_t* map) {
    if (!c->contactsReady) return c->fail(DEME_ERR_NOT_READY, "no contact list");
    const size_t C = c->C;
    const int cu = c->cur;
    if (C) {
        if (idA) DEME_HIP(c, hipMemcpyAsync(idA, c->idA.p, C * 4, hipMemcpyDeviceToHost, c->stream));
        if (idB) DEME_HIP(c, hipMemcpyAsync(idB, c->idB[cu].p, C * 4, hipMemcpyDeviceToHost, c->stream));
        if (type) DEME_HIP(c, hipMemcpyAsync(type, c->ctype[cu].p, C, hipMemcpyDeviceToHost, c->stream));
        if (map) DEME_HIP(c, hipMemcpyAsync(map, c->map.p, C * 4, hipMemcpyDeviceToHost, c->stream));
    }
    DEME_HIP(c, hipStreamSynchronize(c->stream));
    return DEME_OK;
}

// ---------------------------------------------------------------------------------------------------------------
// Dynamics
// ---------------------------------------------------------------------------------------------------------------
// Compiles the specialised force kernel for the context's model, placeholder overrides and constants.
static int jit_rebuild(deme_ctx* c) {
    std::string src, err, log;
    if (!jit_force_source(c->jitModel, c->jitWildcards, c->haveParams ? &c->sp : nullptr, c->jitOverrides, src, err))
        return c->fail(DEME_ERR_INVALID_ARGUMENT, err);
    hipDeviceProp_t prop;
    DEME_HIP(c, hipGetDeviceProperties(&prop, c->device));
    std::vector<char> code;
    if (jit_compile(src, prop.gcnArchName, code, log))
        return c->fail(DEME_ERR_JIT, "hipRTC compilation of the force model failed:\n" + log);
    DEME_HIP(c, c->jit.load(code));
    c->jitStale = false;
    return DEME_OK;
}

// Touch flags for the owner terms: used when lists are detected with a drift margin (most pairs of such a list may
// not touch, and their zero terms are then neither written nor read); without a margin every pair touches at
// detection and every term is written.
static uint8_t* term_flags(deme_ctx* c, bool bSide) {
    if (!(c->sp.beta > 0.f)) return nullptr;
    return bSide ? c->bOn.p : c->aOn.p;
}

// The force kernel of the context: the specialised one when a model or placeholder values were given, else the
// ahead-of-time kernel of the built-in model.
static hipError_t launch_force_ctx(deme_ctx* c, bool full, uint32_t nC, const uint32_t* idB, const uint8_t* ctype,
                                   const OwnerRec* own, const float4* histIn, const uint32_t* map, float4* histOut,
                                   ForceRec* out, SpecGuard guard = SpecGuard()) {
    if (c->jitActive)
        return c->jit.launch(full, c->kp, nC, c->idA.p, idB, ctype, c->sph.p, own, c->moi.p, c->anal.p, c->pose.p,
                             c->mats.p, histIn, map, histOut, c->mirror.p, c->aTerm.p, c->bTerm.p, c->gTerm.p,
                             term_flags(c, false), term_flags(c, true), out, c->stream, guard);
    return launch_force(c->kp, nC, c->idA.p, idB, ctype, c->sph.p, own, c->moi.p, c->anal.p, c->pose.p, c->mats.p,
                        histIn, map, histOut, c->mirror.p, c->aTerm.p, c->bTerm.p, c->gTerm.p, term_flags(c, false),
                        term_flags(c, true), out, c->stream, guard);
}

// The owner update of the context (the specialised kernel when motion strategies or rules were given): owner sums of
// the list at parity cu, then the integration when `integrate`.
static hipError_t launch_owner_ctx(deme_ctx* c, int integrate, int cu, SpecGuard guard = SpecGuard()) {
    if (c->ownerJitActive)
        return c->jitOwner.launch(c->kp, c->nO, c->nClumps, c->nClumps, c->nAnalOwners, c->ownerSphStart.p,
                                  c->rowStart[cu].p, c->bRowStart.p, c->aTerm.p, c->bTerm.p, term_flags(c, false),
                                  term_flags(c, true), c->analAcc.p, c->rules.p, c->own.p, c->own.p, c->acc.p,
                                  integrate, c->clock.p, c->stream, guard);
    return launch_owner_update(c->kp, c->nO, c->nClumps, c->nClumps, c->nAnalOwners, c->ownerSphStart.p,
                               c->rowStart[cu].p, c->bRowStart.p, c->aTerm.p, c->bTerm.p, term_flags(c, false),
                               term_flags(c, true), c->analAcc.p, c->rules.p, c->own.p, c->acc.p, integrate,
                               c->clock.p, c->stream, nullptr, guard);
}

// poseFresh: a detection pass has just computed the plane poses from the same owner state (the same step), so the
// force pass reuses them.
static int force_pass(deme_ctx* c, bool full, bool poseFresh = false) {
    int r = check_ready(c);
    if (r) return r;
    if (c->jitActive && c->jitStale && (r = jit_rebuild(c))) return r;
    if (!c->contactsReady) return c->fail(DEME_ERR_NOT_READY, "contact detection must run before the force pass");
    if (c->needDetect) return c->fail(DEME_ERR_NOT_READY, "clumps were appended: run contact detection first");
    const int cu = c->cur, pv = cu ^ 1;
    const bool remap = c->fresh;
    if (full) DEME_HIP(c, c->frc.ensure(c->C));
    if (!poseFresh)
        DEME_LAUNCH(c, "plane_pose", launch_plane_pose(c->kp, c->anal.p, c->nA, c->own.p, c->pose.p, c->stream));
    DEME_LAUNCH(c, "force",
                launch_force_ctx(c, full, (uint32_t)c->C, c->idB[cu].p, c->ctype[cu].p, c->own.p,
                                 remap ? (c->havePrev ? c->hist[pv].p : c->hist[cu].p) : c->hist[cu].p,
                                 remap ? c->map.p : nullptr, c->hist[cu].p, full ? c->frc.p : nullptr));
    c->fresh = false;
    c->forcesFull = full;
    c->termsReady = true;
    return DEME_OK;
}

int deme_force(deme_ctx* c) { return force_pass(c, true); }

// capturing: enqueued into a graph captured for replay (the clock is then written before each replay instead)
static int collect_and_update(deme_ctx* c, int integrate, bool capturing = false) {
    if (!c->termsReady) return c->fail(DEME_ERR_NOT_READY, "the force pass must run on the current contact list first");
    const int cu = c->cur;
    uint32_t nb = 0;
    DEME_HIP(c, c->analPartials.ensure(std::max<size_t>((size_t)anal_blocks(c->C) * MAX_ANAL_OWNERS * 6, 1)));
    DEME_HIP(c, c->analAcc.ensure(MAX_ANAL_OWNERS * 6, 1.0));
    DEME_LAUNCH(c, "anal_collect",
                launch_anal_collect((uint32_t)c->C, c->idB[cu].p, c->ctype[cu].p, c->gTerm.p, c->anal.p, c->nA,
                                    c->nClumps, c->nAnalOwners, c->analPartials.p, c->analPartials.cap, &nb,
                                    c->stream));
    DEME_LAUNCH(c, "anal_final", launch_anal_final(c->analPartials.p, nb, c->nAnalOwners, c->analAcc.p, c->stream));
    // Motion rules read the time from the device clock; it is written here, before the update, on the stream (a step
    // replayed from a graph gets it from set_step_clock, launched outside the graph)
    if (integrate && c->ownerJitActive && !capturing)
        DEME_LAUNCH(c, "set_clock", launch_set_clock(c->clock.p, c->time, c->stream));
    DEME_LAUNCH(c, integrate ? "owner_update" : "owner_reduce", launch_owner_ctx(c, integrate, cu));
    return DEME_OK;
}

int deme_reduce_owner(deme_ctx* c) {
    int r = check_ready(c);
    if (r) return r;
    if (!c->contactsReady) return c->fail(DEME_ERR_NOT_READY, "no contact list");
    return collect_and_update(c, 0);
}

int deme_integrate(deme_ctx* c) {
    int r = check_ready(c);
    if (r) return r;
    if (!c->contactsReady) return c->fail(DEME_ERR_NOT_READY, "no contact list");
    return collect_and_update(c, 1);
}

int deme_integrate_owners(deme_ctx* c) {
    int r = check_ready(c);
    if (r) return r;
    // the owner kernel reads nothing of the contact list in this mode
    if (c->ownerJitActive)
        DEME_LAUNCH(c, "set_clock", launch_set_clock(c->clock.p, c->time, c->stream));
    DEME_LAUNCH(c, "owner_integrate",
                c->ownerJitActive
                    ? c->jitOwner.launch(c->kp, c->nO, c->nClumps, c->nClumps, c->nAnalOwners, nullptr, nullptr,
                                         nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, c->rules.p, c->own.p,
                                         c->own.p, c->acc.p, 2, c->clock.p, c->stream)
                    : launch_owner_update(c->kp, c->nO, c->nClumps, c->nClumps, c->nAnalOwners, nullptr, nullptr,
                                          nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, c->rules.p, c->own.p,
                                          c->acc.p, 2, c->clock.p, c->stream));
    return DEME_OK;
}

// One non-detection step from the captured graph, capturing it first when anything it depends on changed.
static int step_graph(deme_ctx* c) {
    const bool valid = c->stepExec && c->capturedGen == c->graphGen &&
                       std::memcmp(&c->capturedKp, &c->kp, sizeof(KParams)) == 0;
    if (!valid) {
        hipGraph_t g = nullptr;
        DEME_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int r = force_pass(c, false);
        if (!r) r = collect_and_update(c, 1, true);
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (r) {
            if (e == hipSuccess && g) (void)hipGraphDestroy(g);
            return r;
        }
        DEME_HIP(c, e);
        if (c->stepExec) {
            hipGraphNode_t errNode = nullptr;
            hipGraphExecUpdateResult res;
            if (hipGraphExecUpdate(c->stepExec, g, &errNode, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(c->stepExec);
                c->stepExec = nullptr;
            }
        }
        hipError_t ie = hipSuccess;
        if (!c->stepExec) ie = hipGraphInstantiate(&c->stepExec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        DEME_HIP(c, ie);
        c->capturedGen = c->graphGen;
        c->capturedKp = c->kp;
    }
    // the time of this step for the motion rules of the replayed owner update (not a captured argument)
    if (c->ownerJitActive) DEME_LAUNCH(c, "set_clock", launch_set_clock(c->clock.p, c->time, c->stream));
    DEME_HIP(c, hipGraphLaunch(c->stepExec, c->stream));
    return DEME_OK;
}

// ---- detection intervals captured whole ----
// Everything a captured interval of list parity p baked in (buffer addresses, sizes, grid, kernels); a replay is
// valid only while all of it is unchanged.
static std::vector<const void*> cycle_fingerprint(deme_ctx* c, int p) {
    const int q = p ^ 1;
    auto v = [](uint64_t x) { return reinterpret_cast<const void*>((uintptr_t)x); };
    uint64_t invBits, skinBits, unitBits;
    std::memcpy(&invBits, &c->grid.inv, sizeof(invBits));
    const double skin = c->skin();
    std::memcpy(&skinBits, &skin, sizeof(skinBits));
    std::memcpy(&unitBits, &c->posUnit, sizeof(unitBits));
    return {c->sph.p,        c->own.p,        c->moi.p,         c->anal.p,        c->pose.p,       c->mats.p,
            c->mask.p,       c->rules.p,      c->ownerSphStart.p, c->cellCount.p, c->cellStart.p,  c->cellOf.p,
            c->slotOf.p,     c->posOf.p,      c->aCnt.p,        c->bCnt.p,        c->sortedMeta.p, c->sorted.p,
            c->sortedF.p,    c->planeBits.p,  c->numBins.p,     c->numAnal.p,     c->scanTmp.p,    c->scan2Sums.p,
            c->scanScratch.p, c->survSlab.p,  c->rowSlab.p,     c->mirSlab.p,     c->rowStart[p].p, c->rowStart[q].p,
            c->bRowStart.p,  c->idA.p,        c->idB[p].p,      c->idB[q].p,      c->ctype[p].p,   c->ctype[q].p,
            c->hist[p].p,    c->hist[q].p,    c->map.p,         c->mirror.p,      c->bPart.p,      c->aTerm.p,
            c->bTerm.p,      c->gTerm.p,      c->aOn.p,         c->bOn.p,         c->analPartials.p, c->analAcc.p,
            c->acc.p,        c->clock.p,      c->ctr.p,         c->planeStart.p,
            reinterpret_cast<const void*>(c->jit.full), reinterpret_cast<const void*>(c->jit.terms),
            reinterpret_cast<const void*>(c->jitOwner.fn),
            v(c->slabCap),   v(c->nCells),    v(c->nS),         v(c->nO),         v(c->grid.nx),   v(c->grid.ny),
            v(c->grid.nz),   v(invBits),      v(c->jitActive),  v(c->ownerJitActive), v(c->maskAny),
            c->pos.p,        c->pos0.p,       c->meta.p,        c->candStart.p,   c->candB.p,      v(skinBits),
            c->tStart.p,     c->tC.p,         c->candAcc[0].p,  c->candAcc[1].p,  c->cidx[0].p,    c->cidx[1].p,
            c->candA.p,      v(c->nCand),     c->dispRef.p,     c->tPos.p,        c->tRank.p,      v(unitBits)};
}

// Buffers of the list at parity p (the one the interval's pass writes; the other parity holds the current list and
// is left alone) and the shared per-contact buffers, sized with headroom (room for lists that grow a little between
// passes without a new capture); the guard the interval's kernels check.
static int cycle_capacity(deme_ctx* c, int p, SpecGuard& g) {
    const size_t wantC = std::max<size_t>((size_t)((double)c->C * c->cycleHeadroom) + c->cycleSlack, 1);
    const size_t wantP = std::max<size_t>((size_t)((double)c->nPairs * c->cycleHeadroom) + c->cycleSlack, 1);
    DEME_HIP(c, c->idB[p].ensure(wantC, 1.0));
    DEME_HIP(c, c->ctype[p].ensure(wantC, 1.0));
    DEME_HIP(c, c->hist[p].ensure(wantC, 1.0));
    DEME_HIP(c, c->idA.ensure(wantC, 1.0));
    DEME_HIP(c, c->map.ensure(wantC, 1.0));
    DEME_HIP(c, c->mirror.ensure(wantC, 1.0));
    DEME_HIP(c, c->aTerm.ensure(6 * wantC, 1.0));
    DEME_HIP(c, c->gTerm.ensure(6 * wantC, 1.0));
    DEME_HIP(c, c->aOn.ensure(wantC, 1.0));
    DEME_HIP(c, c->bPart.ensure(wantP, 1.0));
    DEME_HIP(c, c->bTerm.ensure(6 * wantP, 1.0));
    DEME_HIP(c, c->bOn.ensure(wantP, 1.0));
    DEME_HIP(c, c->analPartials.ensure((size_t)std::max(anal_blocks(wantC), 1u) * MAX_ANAL_OWNERS * 6, 1.0));
    DEME_HIP(c, c->analAcc.ensure(MAX_ANAL_OWNERS * 6, 1.0));
    g = SpecGuard();
    g.ctr = c->ctr.p;
    g.slabCap = c->slabCap;
    const uint64_t analCap = (uint64_t)(c->analPartials.cap / (MAX_ANAL_OWNERS * 6)) * ANAL_TILE;
    g.capC = std::min<uint64_t>({c->idB[p].cap, c->ctype[p].cap, c->hist[p].cap, c->idA.cap, c->map.cap,
                                 c->mirror.cap, c->aTerm.cap / 6, c->gTerm.cap / 6, c->aOn.cap, analCap,
                                 0xFFFFFFFFull});
    g.capP = std::min<uint64_t>({c->bPart.cap, c->bTerm.cap / 6, c->bOn.cap});
    return DEME_OK;
}

// The interval's launches with the new list at parity c->cur: the detection pass, the rows and history map, then f
// steps (the first gathers the history through the map), each advancing the device clock by h.
static int enqueue_cycle(deme_ctx* c, int f, const SpecGuard& g, int mode) {
    const int p = c->cur, q = p ^ 1;
    int r;
    if (c->skin() > 0.0) {
        r = enqueue_skin_pass(c, p);
    } else {
        r = enqueue_cell_pass(c, mode);
        if (!r) r = enqueue_pair_pass(c, mode, p);
    }
    if (!r) r = enqueue_rows(c, g, c->skin() > 0.0);
    if (r) return r;
    const uint32_t capC = (uint32_t)g.capC;
    for (int s = 0; s < f; s++) {
        if (s > 0)
            DEME_LAUNCH(c, "plane_pose", launch_plane_pose(c->kp, c->anal.p, c->nA, c->own.p, c->pose.p, c->stream));
        DEME_LAUNCH(c, "force",
                    launch_force_ctx(c, false, capC, c->idB[p].p, c->ctype[p].p, c->own.p,
                                     s == 0 ? c->hist[q].p : c->hist[p].p, s == 0 ? c->map.p : nullptr, c->hist[p].p,
                                     nullptr, g));
        uint32_t nb = 0;
        DEME_LAUNCH(c, "anal_collect",
                    launch_anal_collect(capC, c->idB[p].p, c->ctype[p].p, c->gTerm.p, c->anal.p, c->nA, c->nClumps,
                                        c->nAnalOwners, c->analPartials.p, c->analPartials.cap, &nb, c->stream, g));
        DEME_LAUNCH(c, "anal_final",
                    launch_anal_final(c->analPartials.p, nb, c->nAnalOwners, c->analAcc.p, c->stream, g));
        DEME_LAUNCH(c, "owner_update", launch_owner_ctx(c, 1, p, g));
        DEME_LAUNCH(c, "advance_clock", launch_advance_clock(c->clock.p, c->sp.h, c->stream));
    }
    return DEME_OK;
}

static int capture_cycle(deme_ctx* c, int p, int f) {
    deme_ctx::Cycle& cy = c->cycle[p];
    int mode = 0;
    int r = c->skin() > 0.0 ? skin_buffers(c) : cd_buffers(c, &mode);
    if (r) return r;
    SpecGuard g;
    r = cycle_capacity(c, p, g);
    if (r) return r;
    const int savedCur = c->cur;
    const bool savedPrev = c->havePrev;
    c->cur = p;
    c->havePrev = true;
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        c->cur = savedCur;
        c->havePrev = savedPrev;
        return c->fail_hip(e, "hipStreamBeginCapture");
    }
    r = enqueue_cycle(c, f, g, mode);
    e = hipStreamEndCapture(c->stream, &graph);
    c->cur = savedCur;
    c->havePrev = savedPrev;
    if (r) {
        if (e == hipSuccess && graph) (void)hipGraphDestroy(graph);
        return r;
    }
    DEME_HIP(c, e);
    if (cy.exec) {
        (void)hipGraphExecDestroy(cy.exec);
        cy.exec = nullptr;
    }
    e = hipGraphInstantiate(&cy.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        cy.exec = nullptr;
        return c->fail_hip(e, "hipGraphInstantiate");
    }
    cy.gen = c->graphGen;
    cy.f = f;
    cy.kp = c->kp;
    cy.guard = g;
    cy.ptrs = cycle_fingerprint(c, p);
    c->cycleCaptures++;
    return DEME_OK;
}

// One detection interval of f steps from its captured graph (capturing it first when needed). *done is false when
// the replay's pass failed or outgrew the buffers: then nothing moved and the caller runs the interval directly.
static int run_cycle(deme_ctx* c, int f, bool* done) {
    *done = false;
    const int p = c->cur ^ 1;
    deme_ctx::Cycle& cy = c->cycle[p];
    // a captured candidate-list pass maps the history through the previous list's candidate indices: they must be
    // the current build's
    if (c->skin() > 0.0 && c->cidxGen[c->cur] != c->candBuilds) return DEME_OK;
    const bool valid = cy.exec && cy.gen == c->graphGen && cy.f == f &&
                       std::memcmp(&cy.kp, &c->kp, sizeof(KParams)) == 0 && cy.ptrs == cycle_fingerprint(c, p);
    if (!valid) {
        int r = capture_cycle(c, p, f);
        if (r) return r;
    }
    DEME_LAUNCH(c, "set_clock", launch_set_clock(c->clock.p, c->time, c->stream));
    DEME_HIP(c, hipGraphLaunch(cy.exec, c->stream));
    DEME_HIP(c, hipMemcpyAsync(c->hctr, c->ctr.p, DEV_COUNTERS_HOST_BYTES, hipMemcpyDeviceToHost, c->stream));
    DEME_HIP(c, hipStreamSynchronize(c->stream));
    const DevCounters& h = *c->hctr;
    const SpecGuard& g = cy.guard;
    if (h.err || h.maxOcc > g.slabCap || h.total[1] > g.capC || h.total[2] > g.capP) {
        c->cycleFallbacks++;
        c->cycleBackoff = std::min<uint32_t>(std::max<uint32_t>(2 * c->cycleBackoff, 1), 64);
        c->cycleSkip = c->cycleBackoff;
        if (h.err & ERRBIT_STALE) c->rebuildNext = true;  // the direct interval starts with new candidate lists
        return DEME_OK;
    }
    c->cycleBackoff = 0;
    c->cidxGen[p] = c->skin() > 0.0 ? c->candBuilds : 0;
    end_pass(c, c->skin() > 0.0, max_disp2(h));
    // the bookkeeping of a detection pass and f steps
    c->cur = p;
    c->havePrev = true;
    c->Cprev = c->C;
    c->C = h.total[1];
    c->nPairs = h.total[2];
    c->nGeo = c->C - c->nPairs;
    c->contactsReady = true;
    c->needDetect = false;
    c->fresh = false;
    c->forcesFull = false;
    c->termsReady = true;
    c->cdStage = 0;
    for (int k = 0; k < f; k++) {
        c->time += (double)c->sp.h;
        c->nSteps++;
    }
    c->cycleReplays++;
    *done = true;
    return DEME_OK;
}

int deme_step(deme_ctx* c, int nSteps, int updateFreq) {
    c->adjReady = false;  // recorded adjoint checkpoints no longer describe the system
    int r = check_ready(c);
    if (r) return r;
    const uint64_t f = updateFreq > 0 ? (uint64_t)updateFreq : 1;
    for (int k = 0; k < nSteps; k++) {
        const bool detect = c->nSteps % f == 0 || !c->contactsReady || c->needDetect;
        // a whole interval from its graph: due for detection, a current list to carry the history over, all f steps
        // within this call, and no per-contact records wanted on its last step
        const bool listsReady = c->skin() <= 0.0 || (c->candValid && !c->rebuildNext);
        bool cycle = detect && listsReady && c->useGraphs && c->contactsReady && !c->needDetect && !c->timing &&
                     !c->jitStale && c->stream && c->nS > 0 && (uint64_t)(nSteps - k) >= f && f <= 1000 &&
                     !(c->keepRecords && (uint64_t)(nSteps - k) == f);
        if (cycle && c->cycleSkip > 0) {
            c->cycleSkip--;
            cycle = false;
        }
        if (cycle) {
            bool done = false;
            if ((r = run_cycle(c, (int)f, &done))) return r;
            if (done) {
                k += (int)f - 1;
                continue;
            }
        }
        if (detect) {
            r = deme_cd_run(c);
            if (r) return r;
        }
        // Between detection passes every step launches the same kernels with the same arguments (the history is
        // updated in place; the time the motion rules read is written to the device clock before each replay), so it
        // is captured once per pass and replayed. The step right after a detection pass (history gathered through the
        // map) runs directly.
        const bool records = c->keepRecords && k == nSteps - 1;
        if (!detect && !c->fresh && !records && c->useGraphs && !c->timing && !c->jitStale && c->stream) {
            if ((r = step_graph(c))) return r;
        } else {
            r = force_pass(c, records, detect);
            if (r) return r;
            r = collect_and_update(c, 1);
            if (r) return r;
        }
        c->time += (double)c->sp.h;
        c->nSteps++;
    }
    return DEME_OK;
}

int deme_set_keep_contact_records(deme_ctx* c, int on) {
    c->keepRecords = on != 0;
    return DEME_OK;
}

int deme_set_step_graphs(deme_ctx* c, int on) {
    c->useGraphs = on != 0;
    c->graphGen++;
    return DEME_OK;
}

int deme_set_graph_headroom(deme_ctx* c, double factor, uint64_t slack) {
    if (!(factor >= 1.0) || !(factor <= 16.0))
        return c->fail(DEME_ERR_INVALID_ARGUMENT, "graph headroom factor must lie in [1, 16]");
    c->cycleHeadroom = factor;
    c->cycleSlack = slack;
    c->graphGen++;
    return DEME_OK;
}

int deme_set_skin(deme_ctx* c, double skin) {
    if (!(skin >= -1.0) || !std::isfinite(skin))
        return c->fail(DEME_ERR_INVALID_ARGUMENT, "skin must be a finite length >= 0 (or -1: automatic)");
    c->skinSet = skin < 0.0 ? -1.0 : skin;
    c->skinAuto = 0.0;
    c->skinCostPrev = -1.0;
    c->skinDir = 0;
    c->cellPassesLeft = 0;
    c->cellBackoff = 0;
    c->candValid = false;
    c->graphGen++;
    return DEME_OK;
}

int deme_candidate_stats(deme_ctx* c, double* skin, uint64_t* nCandidates, uint64_t* builds) {
    if (skin) *skin = c->skin();
    if (nCandidates) *nCandidates = c->nCand;
    if (builds) *builds = c->candBuilds;
    return DEME_OK;
}

int deme_set_pair_sweep(deme_ctx* c, int mode, int block) {
    if (mode < -1 || mode > 2) return c->fail(DEME_ERR_INVALID_ARGUMENT, "pair sweep mode must be -1, 0, 1 or 2");
    if (block != 0 && block != 64 && block != 128 && block != 256)
        return c->fail(DEME_ERR_INVALID_ARGUMENT, "pair sweep workgroup size must be 64, 128 or 256 (0: keep)");
    c->pairMode = mode;
    if (block) c->pairBlock = (uint32_t)block;
    c->graphGen++;
    return DEME_OK;
}

int deme_graph_stats(deme_ctx* c, uint64_t out[3]) {
    out[0] = c->cycleCaptures;
    out[1] = c->cycleReplays;
    out[2] = c->cycleFallbacks;
    return DEME_OK;
}

int deme_set_cell_scale(deme_ctx* c, double scale) {
    c->graphGen++;
    if (!(scale >= 1.0) || scale > 8.0) return c->fail(DEME_ERR_INVALID_ARGUMENT, "cell scale must lie in [1, 8]");
    c->cellScale = scale;
    c->nCells = 0;  // re-derive the grid at the next detection pass
    return DEME_OK;
}

int deme_set_beta(deme_ctx* c, float beta) {
    c->graphGen++;
    c->candValid = false;  // positions, radii or the margin changed: candidate lists are rebuilt
    c->adjReady = false;  // recorded adjoint checkpoints no longer describe the system
    if (!(beta >= 0.f) || !std::isfinite(beta)) return c->fail(DEME_ERR_INVALID_ARGUMENT, "beta must be >= 0");
    c->sp.beta = beta;
    c->kp.beta = beta;
    return DEME_OK;
}

// Trial runs of candidate (contact-update interval, cell scale) pairs on the live system; the drift margin of an
// interval f is beta = maxVel * h * f * safety (transferSimParams, APIPrivate.cpp:731-733). Each trial starts with a
// detection pass and runs trialSteps steps (rounded up to a multiple of f). The fastest candidate stays configured.
int deme_autotune(deme_ctx* c, const int* freqs, int nFreqs, const double* cellScales, int nScales, int trialSteps,
                  float maxVel, float safety, int* bestFreq, double* msPerStep) {
    int r = check_ready(c);
    if (r) return r;
    if (nFreqs <= 0 || nScales <= 0 || trialSteps <= 0 || !freqs || !cellScales)
        return c->fail(DEME_ERR_INVALID_ARGUMENT, "need candidates and trial steps");
    if (!(maxVel >= 0.f) || !(safety > 0.f)) return c->fail(DEME_ERR_INVALID_ARGUMENT, "bad velocity or safety");
    if (maxVel == 0.f && c->nO) {
        // measured: the largest owner speed of the current state
        std::vector<OwnerRec> h(c->nO);
        DEME_HIP(c, hipMemcpyAsync(h.data(), c->own.p, c->nO * sizeof(OwnerRec), hipMemcpyDeviceToHost, c->stream));
        DEME_HIP(c, hipStreamSynchronize(c->stream));
        for (const OwnerRec& o : h) maxVel = std::max(maxVel, std::sqrt(o.vx * o.vx + o.vy * o.vy + o.vz * o.vz));
    }
    hipEvent_t e0, e1;
    DEME_HIP(c, hipEventCreate(&e0));
    DEME_HIP(c, hipEventCreate(&e1));
    double best = 1e300;
    int bf = freqs[0];
    double bs = cellScales[0];
    float bb = c->sp.beta;
    for (int i = 0; i < nFreqs; i++) {
        for (int j = 0; j < nScales; j++) {
            const int f = std::max(freqs[i], 1);
            const float beta = f > 1 ? maxVel * c->sp.h * (float)f * safety : 0.f;
            if ((r = deme_set_beta(c, beta)) || (r = deme_set_cell_scale(c, cellScales[j]))) break;
            const int n = ((trialSteps + f - 1) / f) * f;
            c->contactsReady = false;  // the trial starts with a detection pass
            DEME_HIP(c, hipEventRecord(e0, c->stream));
            if ((r = deme_step(c, n, f))) break;
            DEME_HIP(c, hipEventRecord(e1, c->stream));
            DEME_HIP(c, hipEventSynchronize(e1));
            float ms = 0.f;
            DEME_HIP(c, hipEventElapsedTime(&ms, e0, e1));
            const double per = (double)ms / n;
            if (msPerStep) msPerStep[i * nScales + j] = per;
            if (per < best) {
                best = per;
                bf = f;
                bs = cellScales[j];
                bb = beta;
            }
        }
        if (r) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (r) return r;
    deme_set_beta(c, bb);
    deme_set_cell_scale(c, bs);
    c->contactsReady = false;
    *bestFreq = bf;
    return DEME_OK;
}

// Trial runs of the pair sweep's launch configurations (mode x workgroup size) on the live system at update interval
// f, with the context's drift margin and cell scale; each trial runs trialSteps steps (rounded up to a multiple of
// f, so every trial holds the same number of detection passes). The fastest stays configured.
int deme_autotune_sweep(deme_ctx* c, const double* skins, int nSkins, const int* modes, int nModes, const int* blocks,
                        int nBlocks, int trialSteps, int updateFreq, double* bestSkin, int* bestMode, int* bestBlock,
                        double* msPerStep) {
    int r = check_ready(c);
    if (r) return r;
    if (nModes <= 0 || nBlocks <= 0 || trialSteps <= 0 || !modes || !blocks || (skins && nSkins <= 0))
        return c->fail(DEME_ERR_INVALID_ARGUMENT, "need candidates and trial steps");
    const double keepSkin = c->skinSet;
    if (!skins) nSkins = 1;
    const int f = std::max(updateFreq, 1);
    const int n = ((trialSteps + f - 1) / f) * f;
    hipEvent_t e0, e1;
    DEME_HIP(c, hipEventCreate(&e0));
    DEME_HIP(c, hipEventCreate(&e1));
    double best = 1e300;
    double bs = keepSkin;
    int bm = c->pairMode, bb = (int)c->pairBlock;
    for (int k = 0; k < nSkins && !r; k++) {
        if (skins && (r = deme_set_skin(c, skins[k]))) break;
        // with candidate lists the sweep runs only in its build mode: one mode's timings stand for all
        const bool cand = c->skin() > 0.0;
        for (int i = 0; i < nModes && !r; i++) {
            for (int j = 0; j < nBlocks; j++) {
                if (cand && i > 0) {
                    if (msPerStep)
                        msPerStep[((size_t)k * nModes + i) * nBlocks + j] = msPerStep[(size_t)k * nModes * nBlocks + j];
                    continue;
                }
                if ((r = deme_set_pair_sweep(c, modes[i], blocks[j]))) break;
                // a pass with this configuration first (buffers sized, lists built, code loaded), then the timed steps
                if ((r = deme_step(c, f, f))) break;
                DEME_HIP(c, hipEventRecord(e0, c->stream));
                if ((r = deme_step(c, n, f))) break;
                DEME_HIP(c, hipEventRecord(e1, c->stream));
                DEME_HIP(c, hipEventSynchronize(e1));
                float ms = 0.f;
                DEME_HIP(c, hipEventElapsedTime(&ms, e0, e1));
                const double per = (double)ms / n;
                if (msPerStep) msPerStep[((size_t)k * nModes + i) * nBlocks + j] = per;
                if (per < best) {
                    best = per;
                    bs = skins ? skins[k] : keepSkin;
                    bm = modes[i];
                    bb = blocks[j];
                }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (r) return r;
    // the adapted automatic skin is kept when the automatic skin wins its last trial
    if (skins && !(bs < 0.0 && c->skinSet < 0.0) && (r = deme_set_skin(c, bs))) return r;
    r = deme_set_pair_sweep(c, bm, bb);
    if (r) return r;
    if (bestSkin) *bestSkin = bs;
    if (bestMode) *bestMode = bm;
    if (bestBlock) *bestBlock = bb;
    return DEME_OK;
}

// ---------------------------------------------------------------------------------------------------------------
// Adjoint: forward with checkpoints, backward with per-window recomputation.
// ---------------------------------------------------------------------------------------------------------------
static int adj_save(deme_ctx* c, deme_ctx::AdjCk& k, bool full) {
    DEME_HIP(c, k.own.ensure(c->nO, 1.0));
    DEME_HIP(c, hipMemcpyAsync(k.own.p, c->own.p, c->nO * sizeof(OwnerRec), hipMemcpyDeviceToDevice, c->stream));
    k.haveList = c->contactsReady;
    k.C = c->contactsReady ? c->C : 0;
    if (k.haveList) {
        const int cu = c->cur;
        DEME_HIP(c, k.rowStart.ensure(c->nS + 1, 1.0));
        DEME_HIP(c, k.idB.ensure(std::max<uint64_t>(k.C, 1), 1.0));
        DEME_HIP(c, k.ctype.ensure(std::max<uint64_t>(k.C, 1), 1.0));
        DEME_HIP(c, k.hist.ensure(std::max<uint64_t>(k.C, 1), 1.0));
        DEME_HIP(c, hipMemcpyAsync(k.rowStart.p, c->rowStart[cu].p, (c->nS + 1) * 4, hipMemcpyDeviceToDevice, c->stream));
        if (k.C) {
            DEME_HIP(c, hipMemcpyAsync(k.idB.p, c->idB[cu].p, k.C * 4, hipMemcpyDeviceToDevice, c->stream));
            DEME_HIP(c, hipMemcpyAsync(k.ctype.p, c->ctype[cu].p, k.C, hipMemcpyDeviceToDevice, c->stream));
            DEME_HIP(c, hipMemcpyAsync(k.hist.p, c->hist[cu].p, k.C * sizeof(float4), hipMemcpyDeviceToDevice,
                                       c->stream));
        }
        if (full) {
            // idA, mirror slots and mirror rows, so the restored list can be stepped on without a detection pass
            k.nPairs = c->nPairs;
            DEME_HIP(c, k.idA.ensure(std::max<uint64_t>(k.C, 1), 1.0));
            DEME_HIP(c, k.mirror.ensure(std::max<uint64_t>(k.C, 1), 1.0));
            DEME_HIP(c, k.bRowStart.ensure(c->nS + 1, 1.0));
            if (k.C) {
                DEME_HIP(c, hipMemcpyAsync(k.idA.p, c->idA.p, k.C * 4, hipMemcpyDeviceToDevice, c->stream));
                DEME_HIP(c, hipMemcpyAsync(k.mirror.p, c->mirror.p, k.C * 4, hipMemcpyDeviceToDevice, c->stream));
            }
            DEME_HIP(c, hipMemcpyAsync(k.bRowStart.p, c->bRowStart.p, (c->nS + 1) * 4, hipMemcpyDeviceToDevice,
                                       c->stream));
        }
    }
    if (full) {
        DEME_HIP(c, k.acc.ensure((size_t)6 * c->nO, 1.0));
        DEME_HIP(c, hipMemcpyAsync(k.acc.p, c->acc.p, (size_t)6 * c->nO * sizeof(float), hipMemcpyDeviceToDevice,
                                   c->stream));
    }
    k.full = full;
    k.time = c->time;
    k.nSteps = c->nSteps;
    return DEME_OK;
}

static int adj_restore(deme_ctx* c, const deme_ctx::AdjCk& k) {
    c->graphGen++;
    c->candValid = false;  // positions, radii or the margin changed: candidate lists are rebuilt
    DEME_HIP(c, hipMemcpyAsync(c->own.p, k.own.p, c->nO * sizeof(OwnerRec), hipMemcpyDeviceToDevice, c->stream));
    const int cu = c->cur;
    if (k.haveList) {
        DEME_HIP(c, c->idB[cu].ensure(std::max<uint64_t>(k.C, 1)));
        DEME_HIP(c, c->ctype[cu].ensure(std::max<uint64_t>(k.C, 1)));
        DEME_HIP(c, c->hist[cu].ensure(std::max<uint64_t>(k.C, 1)));
        DEME_HIP(c, hipMemcpyAsync(c->rowStart[cu].p, k.rowStart.p, (c->nS + 1) * 4, hipMemcpyDeviceToDevice, c->stream));
        if (k.C) {
            DEME_HIP(c, hipMemcpyAsync(c->idB[cu].p, k.idB.p, k.C * 4, hipMemcpyDeviceToDevice, c->stream));
            DEME_HIP(c, hipMemcpyAsync(c->ctype[cu].p, k.ctype.p, k.C, hipMemcpyDeviceToDevice, c->stream));
            DEME_HIP(c, hipMemcpyAsync(c->hist[cu].p, k.hist.p, k.C * sizeof(float4), hipMemcpyDeviceToDevice,
                                       c->stream));
        }
        c->C = k.C;
        if (k.full) {
            c->nPairs = k.nPairs;
            DEME_HIP(c, c->idA.ensure(std::max<uint64_t>(k.C, 1)));
            DEME_HIP(c, c->mirror.ensure(std::max<uint64_t>(k.C, 1)));
            DEME_HIP(c, c->bTerm.ensure(std::max<uint64_t>((uint64_t)6 * k.nPairs, 1)));
            DEME_HIP(c, c->aTerm.ensure(std::max<uint64_t>((uint64_t)6 * k.C, 1)));
            DEME_HIP(c, c->bOn.ensure(std::max<uint64_t>(k.nPairs, 1)));
            DEME_HIP(c, c->aOn.ensure(std::max<uint64_t>(k.C, 1)));
            DEME_HIP(c, c->gTerm.ensure(std::max<uint64_t>((uint64_t)6 * k.C, 1)));
            if (k.C) {
                DEME_HIP(c, hipMemcpyAsync(c->idA.p, k.idA.p, k.C * 4, hipMemcpyDeviceToDevice, c->stream));
                DEME_HIP(c, hipMemcpyAsync(c->mirror.p, k.mirror.p, k.C * 4, hipMemcpyDeviceToDevice, c->stream));
            }
            DEME_HIP(c, hipMemcpyAsync(c->bRowStart.p, k.bRowStart.p, (c->nS + 1) * 4, hipMemcpyDeviceToDevice,
                                       c->stream));
        }
    }
    if (k.full)
        DEME_HIP(c, hipMemcpyAsync(c->acc.p, k.acc.p, (size_t)6 * c->nO * sizeof(float), hipMemcpyDeviceToDevice,
                                   c->stream));
    c->contactsReady = k.haveList;
    c->havePrev = false;
    c->fresh = false;
    c->termsReady = false;
    c->time = k.time;
    c->nSteps = k.nSteps;
    return DEME_OK;
}

// One window: a detection pass, then n steps of force -> reduce -> integrate. With record, the steps run on per-step
// snapshot buffers instead of in place: owner records adjOwn[k] (before step k) -> adjOwn[k + 1], history
// adjHist[k] -> adjHist[k + 1], accelerations into adjAcc[k]; the reverse sweep reads them back. The analytical
// owners are fixed walls on this path, so their force sums (which no sphere reads) and poses are not recomputed per
// step.
static int adj_window(deme_ctx* c, int n, bool record) {
    int r = deme_cd_run(c);
    if (r) return r;
    if (!record) {
        for (int k = 0; k < n; k++) {
            if ((r = force_pass(c, false))) return r;
            if ((r = collect_and_update(c, 1))) return r;
            c->time += (double)c->sp.h;
            c->nSteps++;
        }
        return DEME_OK;
    }
    if ((int)c->adjOwn.size() < n + 1) {
        c->adjOwn.resize(n + 1);
        c->adjHist.resize(n + 1);
        c->adjAcc.resize(n + 1);
    }
    const uint64_t C = c->C;
    for (int k = 0; k <= n; k++) {
        DEME_HIP(c, c->adjOwn[k].ensure(c->nO));
        DEME_HIP(c, c->adjHist[k].ensure(std::max<uint64_t>(C, 1)));
        DEME_HIP(c, c->adjAcc[k].ensure((size_t)6 * c->nO));
    }
    DEME_HIP(c, c->analAcc.ensure(MAX_ANAL_OWNERS * 6, 1.0));
    const int cu = c->cur, pv = cu ^ 1;
    DEME_LAUNCH(c, "plane_pose", launch_plane_pose(c->kp, c->anal.p, c->nA, c->own.p, c->pose.p, c->stream));
    DEME_LAUNCH(c, "adj_gather", launch_adj_gather_hist((uint32_t)C, c->havePrev ? c->map.p : nullptr,
                                                         (uint32_t)c->Cprev, c->hist[pv].p, c->adjHist[0].p,
                                                         c->stream));
    DEME_HIP(c, hipMemcpyAsync(c->adjOwn[0].p, c->own.p, c->nO * sizeof(OwnerRec), hipMemcpyDeviceToDevice,
                               c->stream));
    for (int k = 0; k < n; k++) {
        DEME_LAUNCH(c, "force", launch_force_ctx(c, false, (uint32_t)C, c->idB[cu].p, c->ctype[cu].p, c->adjOwn[k].p,
                                                 c->adjHist[k].p, nullptr, c->adjHist[k + 1].p, nullptr));
        DEME_LAUNCH(c, "owner_update",
                    launch_owner_update(c->kp, c->nO, c->nClumps, c->nClumps, c->nAnalOwners, c->ownerSphStart.p,
                                        c->rowStart[cu].p, c->bRowStart.p, c->aTerm.p, c->bTerm.p,
                                        term_flags(c, false), term_flags(c, true),
                                          c->analAcc.p,
                                        c->rules.p, c->adjOwn[k].p, c->adjAcc[k].p, 1, c->clock.p, c->stream,
                                        c->adjOwn[k + 1].p));
        c->time += (double)c->sp.h;
        c->nSteps++;
    }
    c->fresh = false;
    c->termsReady = false;
    return DEME_OK;
}

static int adj_check(deme_ctx* c) {
    if (c->jitCustom)
        return c->fail(DEME_ERR_UNSUPPORTED, "the adjoint differentiates the built-in frictional Hertz-Mindlin model");
    if (c->ownerJitActive)
        return c->fail(DEME_ERR_UNSUPPORTED, "the adjoint has no expression prescriptions or family change rules");
    if (!c->singleSpheres)
        return c->fail(DEME_ERR_UNSUPPORTED, "the adjoint handles single-sphere owners (spheres as their own owners)");
    if (c->nM * 4 > 256) return c->fail(DEME_ERR_UNSUPPORTED, "the adjoint handles at most 64 materials");
    for (const deme_material& m : c->hostMats)
        if (m.Crr != 0.f) return c->fail(DEME_ERR_UNSUPPORTED, "the adjoint has no rolling-resistance term (Crr = 0)");
    if (c->sp.integrator != DEME_INTEGRATOR_EXTENDED_TAYLOR)
        return c->fail(DEME_ERR_UNSUPPORTED, "the adjoint differentiates the extended-Taylor integrator");
    return DEME_OK;
}

int deme_adjoint_forward(deme_ctx* c, int nSteps, int window) {
    int r = check_ready(c);
    if (r) return r;
    if ((r = adj_check(c))) return r;
    if (nSteps <= 0 || window <= 0) return c->fail(DEME_ERR_INVALID_ARGUMENT, "need positive steps and window");
    const int nWin = (nSteps + window - 1) / window;
    c->adjCk.resize(nWin + 1);
    for (int w = 0; w < nWin; w++) {
        if ((r = adj_save(c, c->adjCk[w], false))) return r;
        if ((r = adj_window(c, std::min(window, nSteps - w * window), false))) return r;
    }
    if ((r = adj_save(c, c->adjCk[nWin], true))) return r;
    DEME_HIP(c, hipStreamSynchronize(c->stream));
    c->adjWindow = window;
    c->adjStepsTotal = nSteps;
    c->adjReady = true;
    return DEME_OK;
}

int deme_adjoint_backward(deme_ctx* c, const float* dLdPos, const float* dLdVel, float* gradVel0, float* gradMat) {
    if (!c->adjReady) return c->fail(DEME_ERR_NOT_READY, "deme_adjoint_forward must run first");
    int r;
    const uint32_t nO = c->nO;
    const int window = c->adjWindow, nSteps = c->adjStepsTotal;
    const int nWin = (nSteps + window - 1) / window;
    const uint32_t nMatV = c->nM * 4;
    DEME_HIP(c, c->lam.ensure((size_t)9 * nO, 1.0));
    DEME_HIP(c, c->lAcc.ensure((size_t)6 * nO, 1.0));
    DEME_HIP(c, c->matGrad.ensure(std::max<uint32_t>(nMatV, 1), 1.0));
    DEME_HIP(c, hipMemsetAsync(c->lam.p, 0, (size_t)9 * nO * sizeof(float), c->stream));
    DEME_HIP(c, hipMemsetAsync(c->matGrad.p, 0, nMatV * sizeof(float), c->stream));
    if (dLdPos) DEME_HIP(c, hipMemcpyAsync(c->lam.p, dLdPos, (size_t)3 * nO * sizeof(float), hipMemcpyHostToDevice,
                                           c->stream));
    if (dLdVel) DEME_HIP(c, hipMemcpyAsync(c->lam.p + (size_t)3 * nO, dLdVel, (size_t)3 * nO * sizeof(float),
                                           hipMemcpyHostToDevice, c->stream));
    // the last step's history out feeds nothing
    int cur = 0;
    bool lDtSet = false;
    for (int w = nWin - 1; w >= 0; w--) {
        const int n = std::min(window, nSteps - w * window);
        if ((r = adj_restore(c, c->adjCk[w]))) return r;
        if ((r = adj_window(c, n, true))) return r;
        const uint64_t C = c->C;
        const int cu = c->cur;
        DEME_HIP(c, c->lDt[0].ensure(std::max<uint64_t>(C, 1)));
        DEME_HIP(c, c->lDt[1].ensure(std::max<uint64_t>(C, 1)));
        DEME_HIP(c, c->adjA.ensure(std::max<uint64_t>((uint64_t)9 * C, 1)));
        DEME_HIP(c, c->adjB.ensure(std::max<uint64_t>((uint64_t)9 * c->nPairs, 1)));
        const uint32_t nb = adj_force_blocks((uint32_t)C);
        DEME_HIP(c, c->matPart.ensure(std::max<uint64_t>((uint64_t)nb * nMatV, 1)));
        if (!lDtSet) {
            DEME_LAUNCH(c, "adj_fill", launch_fill_f4(c->lDt[cur].p, C, c->stream));
            lDtSet = true;
        }
        // the window's last step through the integration; then per step the contacts, and the owners' contact sums
        // fused with the integration of the step before
        DEME_LAUNCH(c, "adj_owner_pre", launch_adj_owner_pre(nO, c->nClumps, c->sp.h, c->adjOwn[n - 1].p,
                                                             c->adjAcc[n - 1].p, c->lam.p, c->lAcc.p, c->stream));
        for (int k = n - 1; k >= 0; k--) {
            DEME_LAUNCH(c, "adj_force",
                        launch_adj_force(c->kp, (uint32_t)C, c->nClumps, c->idA.p, c->idB[cu].p, c->ctype[cu].p,
                                         c->sph.p, c->adjOwn[k].p, c->moi.p, c->anal.p, c->pose.p, c->mats.p, c->nM,
                                         c->adjHist[k].p, c->mirror.p, nO, c->lAcc.p, c->lDt[cur].p, c->adjA.p,
                                         c->adjB.p, c->lDt[cur ^ 1].p, c->matPart.p, k != n - 1, c->stream));
            DEME_LAUNCH(c, "adj_owner_post",
                        launch_adj_owner_post(nO, c->nClumps, c->ownerSphStart.p, c->rowStart[cu].p, c->bRowStart.p,
                                              c->adjA.p, c->adjB.p, c->lam.p, c->sp.h,
                                              k > 0 ? c->adjOwn[k - 1].p : nullptr,
                                              k > 0 ? c->adjAcc[k - 1].p : nullptr, c->lAcc.p, c->stream));
            cur ^= 1;  // the history-in adjoint of step k is the history-out adjoint of step k - 1
        }
        DEME_LAUNCH(c, "adj_mat", launch_adj_mat_final(c->matPart.p, C ? nb : 0, nMatV, c->matGrad.p, c->stream));
        if (w > 0) {
            // back through the gather of the window's first step onto the previous window's final list
            const uint64_t Cprev = c->adjCk[w].C;
            DEME_HIP(c, c->lDtPrev.ensure(std::max<uint64_t>(Cprev, 1)));
            DEME_LAUNCH(c, "adj_scatter", launch_adj_scatter_hist((uint32_t)C, c->havePrev ? c->map.p : nullptr,
                                                                  (uint32_t)Cprev, c->lDt[cur].p, c->lDtPrev.p,
                                                                  c->stream));
            DEME_HIP(c, c->lDt[cur].ensure(std::max<uint64_t>(Cprev, 1)));
            DEME_HIP(c, c->lDt[cur ^ 1].ensure(std::max<uint64_t>(Cprev, 1)));
            if (Cprev)
                DEME_HIP(c, hipMemcpyAsync(c->lDt[cur].p, c->lDtPrev.p, Cprev * sizeof(float4),
                                           hipMemcpyDeviceToDevice, c->stream));
        }
    }
    if (gradVel0)
        DEME_HIP(c, hipMemcpyAsync(gradVel0, c->lam.p + (size_t)3 * nO, (size_t)3 * nO * sizeof(float),
                                   hipMemcpyDeviceToHost, c->stream));
    if (gradMat && nMatV)
        DEME_HIP(c, hipMemcpyAsync(gradMat, c->matGrad.p, nMatV * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    // leave the system where the forward run ended
    if ((r = adj_restore(c, c->adjCk[nWin]))) return r;
    DEME_HIP(c, hipStreamSynchronize(c->stream));
    return DEME_OK;
}

// ---------------------------------------------------------------------------------------------------------------
// Force model and runtime specialisation.
// ---------------------------------------------------------------------------------------------------------------
const char* deme_builtin_force_model(int which) {
    return (which == 0 || which == 1) ? jit_builtin_model(which) : nullptr;
}

int deme_set_force_model(deme_ctx* c, const char* model, const char* const* wildcards, int nWildcards) {
    c->graphGen++;
    c->adjReady = false;
    if (!model) {
        c->jit.release();
        c->jitActive = c->jitStale = c->jitCustom = false;
        c->jitOverrides.clear();
        return DEME_OK;
    }
    if (nWildcards < 0 || (nWildcards > 0 && !wildcards))
        return c->fail(DEME_ERR_INVALID_ARGUMENT, "wildcard names missing");
    std::vector<std::string> w;
    for (int k = 0; k < nWildcards; k++) {
        if (!wildcards[k]) return c->fail(DEME_ERR_INVALID_ARGUMENT, "null wildcard name");
        w.emplace_back(wildcards[k]);
    }
    c->jitModel = model;
    c->jitWildcards = w;
    c->jitOverrides.clear();
    c->jitActive = true;
    c->jitStale = true;
    c->jitCustom = true;
    return jit_rebuild(c);
}

int deme_jit_build(deme_ctx* c, const char* const* keys, const char* const* values, int n) {
    c->graphGen++;
    c->adjReady = false;
    if (n < 0 || (n > 0 && (!keys || !values))) return c->fail(DEME_ERR_INVALID_ARGUMENT, "keys/values missing");
    std::vector<std::pair<std::string, std::string>> ov, own;
    for (int k = 0; k < n; k++) {
        if (!keys[k] || !values[k]) return c->fail(DEME_ERR_INVALID_ARGUMENT, "null placeholder key or value");
        (jit_is_owner_key(keys[k]) ? own : ov).emplace_back(keys[k], values[k]);
    }
    if (!own.empty()) {
        // the owner update kernel's placeholders: specialise that kernel
        std::string log;
        hipDeviceProp_t prop;
        DEME_HIP(c, hipGetDeviceProperties(&prop, c->device));
        std::vector<char> code;
        if (jit_compile(jit_owner_source(own), prop.gcnArchName, code, log))
            return c->fail(DEME_ERR_JIT, "hipRTC compilation of the motion rules failed:\n" + log);
        DEME_HIP(c, c->jitOwner.load(code));
        c->ownerValues = own;
        c->ownerJitActive = true;
        if (ov.empty()) return DEME_OK;
    }
    if (!c->jitActive) {
        c->jitModel = jit_builtin_model(0);
        c->jitWildcards = {"delta_tan_x", "delta_tan_y", "delta_tan_z", "delta_time"};
        c->jitCustom = false;
    }
    for (const auto& kv : ov)
        if (kv.first == "_DEMForceModel_" || kv.first.rfind("_forceModel", 0) == 0) c->jitCustom = true;
    c->jitOverrides = ov;
    c->jitActive = true;
    c->jitStale = true;
    return jit_rebuild(c);
}

int deme_clear_motion_rules(deme_ctx* c) {
    c->graphGen++;
    c->adjReady = false;
    c->jitOwner.release();
    c->ownerJitActive = false;
    c->ownerValues.clear();
    return DEME_OK;
}

int deme_change_family(deme_ctx* c, int from, int to) {
    if (from < 0 || from > 255 || to < 0 || to > 255)
        return c->fail(DEME_ERR_INVALID_ARGUMENT, "family numbers lie in 0..255");
    if (!c->haveState) return c->fail(DEME_ERR_NOT_READY, "no owner state");
    c->adjReady = false;
    DEME_LAUNCH(c, "change_family", launch_change_family(c->own.p, c->nO, (uint8_t)from, (uint8_t)to, c->stream));
    return DEME_OK;
}

int deme_download_forces(deme_ctx* c, float* F, float* T, float* cA, float* cB) {
    if (!c->contactsReady) return c->fail(DEME_ERR_NOT_READY, "no contact list");
    if (!c->forcesFull)
        return c->fail(DEME_ERR_NOT_READY, "contact records are kept by deme_force only (deme_step stores owner terms)");
    std::vector<ForceRec> h(c->C);
    if (c->C) DEME_HIP(c, hipMemcpyAsync(h.data(), c->frc.p, c->C * sizeof(ForceRec), hipMemcpyDeviceToHost, c->stream));
    DEME_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < c->C; i++) {
        const ForceRec& f = h[i];
        if (F) { F[3 * i] = f.fx; F[3 * i + 1] = f.fy; F[3 * i + 2] = f.fz; }
        if (T) { T[3 * i] = f.tx; T[3 * i + 1] = f.ty; T[3 * i + 2] = f.tz; }
        if (cA) { cA[3 * i] = f.ax; cA[3 * i + 1] = f.ay; cA[3 * i + 2] = f.az; }
        if (cB) { cB[3 * i] = f.bx; cB[3 * i + 1] = f.by; cB[3 * i + 2] = f.bz; }
    }
    return DEME_OK;
}

int deme_download_wildcards(deme_ctx* c, float* w) {
    if (!c->contactsReady) return c->fail(DEME_ERR_NOT_READY, "no contact list");
    const size_t C = c->C;
    std::vector<float4> h(C);
    if (c->fresh) {
        // Between a detection pass and the next force pass the history still lives in the previous order; the force
        // kernel gathers it through the persistent-contact map. Apply the same gather here.
        std::vector<uint32_t> m(C);
        std::vector<float4> prev(c->havePrev ? c->Cprev : 0);
        if (C) DEME_HIP(c, hipMemcpyAsync(m.data(), c->map.p, C * 4, hipMemcpyDeviceToHost, c->stream));
        if (!prev.empty())
            DEME_HIP(c, hipMemcpyAsync(prev.data(), c->hist[c->cur ^ 1].p, prev.size() * sizeof(float4),
                                       hipMemcpyDeviceToHost, c->stream));
        DEME_HIP(c, hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < C; i++)
            h[i] = (m[i] == DEME_NULL_MAPPING_PARTNER || m[i] >= prev.size()) ? make_float4(0.f, 0.f, 0.f, 0.f)
                                                                             : prev[m[i]];
    } else {
        if (C)
            DEME_HIP(c, hipMemcpyAsync(h.data(), c->hist[c->cur].p, C * sizeof(float4), hipMemcpyDeviceToHost,
                                       c->stream));
        DEME_HIP(c, hipStreamSynchronize(c->stream));
    }
    for (size_t i = 0; i < C; i++) {
        w[i] = h[i].x;
        w[C + i] = h[i].y;
        w[2 * C + i] = h[i].z;
        w[3 * C + i] = h[i].w;
    }
    return DEME_OK;
}

int deme_upload_wildcards(deme_ctx* c, const float* w) {
    c->graphGen++;
    if (!c->contactsReady) return c->fail(DEME_ERR_NOT_READY, "no contact list");
    const size_t C = c->C;
    std::vector<float4> h(C);
    for (size_t i = 0; i < C; i++) h[i] = make_float4(w[i], w[C + i], w[2 * C + i], w[3 * C + i]);
    if (C) DEME_HIP(c, hipMemcpyAsync(c->hist[c->cur].p, h.data(), C * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    DEME_HIP(c, hipStreamSynchronize(c->stream));
    c->fresh = false;  // the uploaded history is already in the current order
    return DEME_OK;
}

int deme_get_counts(deme_ctx* c, deme_counts* out) {
    std::memset(out, 0, sizeof(*out));
    out->nSpheres = c->nS;
    out->nOwners = c->nO;
    out->nAnal = c->nA;
    out->nBinSphereTouches = c->P;
    out->nActiveBins = c->nActive;
    out->nContacts = c->C;
    out->nPrevContacts = c->Cprev;
    out->nSphereGeoContacts = c->nGeo;
    out->maxSpheresInBin = c->maxOcc;
    out->nSteps = c->nSteps;
    out->time = c->time;
    out->nCells = c->nCells;
    return DEME_OK;
}

// ---------------------------------------------------------------------------------------------------------------
// Mid-simulation changes: appended clumps, clump sizes
// ---------------------------------------------------------------------------------------------------------------
int deme_append_clumps(deme_ctx* c, uint32_t nNewSph, const uint32_t* sphOwner, const float* sphRadius,
                       const float* sphRelPos, const uint8_t* sphMat, uint32_t nNew, const float* ownerMass,
                       const float* ownerMoi, const deme_owner_state* s) {
    c->candValid = false;  // positions, radii or the margin changed: candidate lists are rebuilt
    int r = check_ready(c);
    if (r) return r;
    if (c->nScenes != 1) return c->fail(DEME_ERR_UNSUPPORTED, "clumps cannot be appended to a batch of scenes");
    if (nNew == 0) return DEME_OK;
    if (!s || !s->voxelID || !s->locX || !s->locY || !s->locZ || !ownerMass || !ownerMoi ||
        (nNewSph && (!sphOwner || !sphRadius || !sphRelPos || !sphMat)))
        return c->fail(DEME_ERR_INVALID_ARGUMENT, "appended clumps need geometry, mass, MOI and positions");
    const uint32_t nS0 = c->nS, nO0 = c->nO, nC0 = c->nClumps, nAO = c->nO - c->nClumps;
    if ((uint64_t)nS0 + nNewSph > 0xFFFFFFF0ull || (uint64_t)nO0 + nNew > 0xFFFFFFF0ull)
        return c->fail(DEME_ERR_CAPACITY, "too many spheres or owners");
    // new component spheres follow the existing ones; their owners follow the existing clumps
    std::vector<SphereRec> hs(nNewSph);
    std::vector<uint32_t> cnt(nNew, 0);
    uint32_t maxMat = c->maxMatSph;
    float maxR = c->maxRadius;
    for (uint32_t k = 0; k < nNewSph; k++) {
        if (sphOwner[k] >= nNew) return c->fail(DEME_ERR_INVALID_ARGUMENT, "sphere owner is not an appended clump");
        if (k > 0 && sphOwner[k] < sphOwner[k - 1])
            return c->fail(DEME_ERR_INVALID_ARGUMENT, "spheres must be grouped by owner in ascending order");
        if (!std::isfinite(sphRadius[k]) || sphRadius[k] < 0.f)
            return c->fail(DEME_ERR_INVALID_ARGUMENT, "sphere radii must be finite and non-negative");
        if (sphMat[k] >= c->nM) return c->fail(DEME_ERR_INVALID_ARGUMENT, "a material index exceeds the material table");
        SphereRec& h = hs[k];
        h.owner = nC0 + sphOwner[k];
        h.r = sphRadius[k];
        h.rx = sphRelPos[3 * k];
        h.ry = sphRelPos[3 * k + 1];
        h.rz = sphRelPos[3 * k + 2];
        h.mat = sphMat[k];
        h.pad0 = h.pad1 = 0;
        cnt[sphOwner[k]]++;
        maxMat = std::max<uint32_t>(maxMat, sphMat[k]);
        maxR = std::max(maxR, sphRadius[k]);
    }
    const uint32_t nS = nS0 + nNewSph, nO = nO0 + nNew, nC = nC0 + nNew;
    // owner records: [old clumps][new clumps][external objects moved up by nNew]
    std::vector<OwnerRec> tail(nNew + nAO);
    std::vector<float4> tailMoi(nNew + nAO);
    std::vector<float> accOld((size_t)6 * nO0);
    if (nAO) {
        DEME_HIP(c, hipMemcpyAsync(tail.data() + nNew, c->own.p + nC0, nAO * sizeof(OwnerRec), hipMemcpyDeviceToHost,
                                   c->stream));
        DEME_HIP(c, hipMemcpyAsync(tailMoi.data() + nNew, c->moi.p + nC0, nAO * sizeof(float4), hipMemcpyDeviceToHost,
                                   c->stream));
    }
    if (nO0) DEME_HIP(c, hipMemcpyAsync(accOld.data(), c->acc.p, accOld.size() * 4, hipMemcpyDeviceToHost, c->stream));
    std::vector<deme_anal_obj> anal(c->nA);
    if (c->nA)
        DEME_HIP(c, hipMemcpyAsync(anal.data(), c->anal.p, c->nA * sizeof(deme_anal_obj), hipMemcpyDeviceToHost,
                                   c->stream));
    DEME_HIP(c, hipStreamSynchronize(c->stream));
    for (uint32_t k = 0; k < nNew; k++) {
        OwnerRec& o = tail[k];
        std::memset(&o, 0, sizeof(o));
        o.voxel = s->voxelID[k];
        o.lx = s->locX[k];
        o.ly = s->locY[k];
        o.lz = s->locZ[k];
        o.fam = s->family ? s->family[k] : 0;
        o.qw = s->oriQw ? s->oriQw[k] : 1.f;
        o.qx = s->oriQx ? s->oriQx[k] : 0.f;
        o.qy = s->oriQy ? s->oriQy[k] : 0.f;
        o.qz = s->oriQz ? s->oriQz[k] : 0.f;
        o.vx = s->vX ? s->vX[k] : 0.f;
        o.vy = s->vY ? s->vY[k] : 0.f;
        o.vz = s->vZ ? s->vZ[k] : 0.f;
        o.wx = s->omgX ? s->omgX[k] : 0.f;
        o.wy = s->omgY ? s->omgY[k] : 0.f;
        o.wz = s->omgZ ? s->omgZ[k] : 0.f;
        o.mass = ownerMass[k];
        o.scene = 0;
        tailMoi[k] = make_float4(ownerMoi[3 * k], ownerMoi[3 * k + 1], ownerMoi[3 * k + 2], 0.f);
    }
    // accelerations are kept component-major over all owners: re-laid for the new owner count
    std::vector<float> accNew((size_t)6 * nO, 0.f);
    for (int q = 0; q < 6; q++)
        for (uint32_t o = 0; o < nO0; o++)
            accNew[(size_t)q * nO + (o < nC0 ? o : o + nNew)] = accOld[(size_t)q * nO0 + o];
    std::vector<uint32_t> start(nNew);
    for (uint32_t k = 0, at = nS0; k < nNew; k++) {
        at += cnt[k];
        start[k] = at;  // ownerSphStart[nC0 + 1 + k]
    }
    for (auto& a : anal) a.owner += nNew;

    DEME_HIP(c, c->sph.grow_keep(nS, nS0, c->stream));
    DEME_HIP(c, c->own.grow_keep(nO, nC0, c->stream));
    DEME_HIP(c, c->moi.grow_keep(nO, nC0, c->stream));
    DEME_HIP(c, c->ownerSphStart.grow_keep(nC + 1, nC0 + 1, c->stream));
    DEME_HIP(c, c->rowStart[c->cur].grow_keep(nS + 1, c->contactsReady ? nS0 + 1 : 0, c->stream));
    DEME_HIP(c, c->rowStart[c->cur ^ 1].ensure(nS + 1));
    DEME_HIP(c, c->acc.ensure((size_t)6 * nO));
    DEME_HIP(c, c->numBins.ensure(nS));
    DEME_HIP(c, c->numAnal.ensure(nS));
    DEME_HIP(c, c->rec.ensure(nS));
    DEME_HIP(c, c->rowCount.ensure(nS));
    DEME_HIP(c, c->bCnt.ensure(nS + 1));
    DEME_HIP(c, c->bRowStart.ensure(nS + 1));
    DEME_HIP(c, c->scanScratch.ensure(scan_scratch_elems(std::max<uint64_t>({c->nBins, (uint64_t)nS, 1u << 22}))));
    if (nNewSph)
        DEME_HIP(c, hipMemcpyAsync(c->sph.p + nS0, hs.data(), nNewSph * sizeof(SphereRec), hipMemcpyHostToDevice,
                                   c->stream));
    DEME_HIP(c, hipMemcpyAsync(c->own.p + nC0, tail.data(), tail.size() * sizeof(OwnerRec), hipMemcpyHostToDevice,
                               c->stream));
    DEME_HIP(c, hipMemcpyAsync(c->moi.p + nC0, tailMoi.data(), tailMoi.size() * sizeof(float4), hipMemcpyHostToDevice,
                               c->stream));
    DEME_HIP(c, hipMemcpyAsync(c->ownerSphStart.p + nC0 + 1, start.data(), nNew * 4, hipMemcpyHostToDevice, c->stream));
    DEME_HIP(c, hipMemcpyAsync(c->acc.p, accNew.data(), accNew.size() * 4, hipMemcpyHostToDevice, c->stream));
    if (c->nA)
        DEME_HIP(c, hipMemcpyAsync(c->anal.p, anal.data(), c->nA * sizeof(deme_anal_obj), hipMemcpyHostToDevice,
                                   c->stream));
    DEME_HIP(c, hipMemsetAsync(c->rowCount.p, 0, (size_t)nS * 4, c->stream));
    DEME_HIP(c, hipMemsetAsync(c->bCnt.p, 0, (size_t)(nS + 1) * 4, c->stream));
    // the current list becomes the previous one at the next detection pass, which maps histories by sphere id: the
    // appended spheres have empty rows in it
    if (c->contactsReady)
        DEME_LAUNCH(c, "fill", launch_fill_u32(c->rowStart[c->cur].p + nS0 + 1, nNewSph, (uint32_t)c->C, c->stream));
    DEME_HIP(c, hipStreamSynchronize(c->stream));

    c->hostMass.insert(c->hostMass.begin() + nC0, ownerMass, ownerMass + nNew);
    c->hostScene.insert(c->hostScene.begin() + nC0, nNew, 0u);
    for (auto& o : c->hostAnalOwner) o += nNew;
    c->hostRadius.insert(c->hostRadius.end(), sphRadius, sphRadius + nNewSph);
    c->meanRadius = mean_radius(c->hostRadius);
    bool single = c->singleSpheres && nNewSph == nNew;
    for (uint32_t k = 0; k < nNewSph && single; k++)
        single = sphOwner[k] == k && sphRelPos[3 * k] == 0.f && sphRelPos[3 * k + 1] == 0.f &&
                 sphRelPos[3 * k + 2] == 0.f;
    c->singleSpheres = single;
    c->nS = nS;
    c->kp.nS = nS;
    c->nO = nO;
    c->nClumps = nC;
    c->maxMatSph = maxMat;
    c->maxMat = std::max(c->maxMatSph, c->maxMatAnal);
    if (maxR > c->maxRadius) c->nCells = 0;  // re-derive the cell grid for the larger spheres
    c->maxRadius = maxR;
    c->scenesDirty = true;
    c->haveVolumes = false;
    c->needDetect = true;
    c->termsReady = c->forcesFull = false;
    c->adjReady = false;
    c->graphGen++;
    return DEME_OK;
}

int deme_change_clump_sizes(deme_ctx* c, uint32_t n, const uint32_t* ids, const float* factors) {
    c->candValid = false;  // positions, radii or the margin changed: candidate lists are rebuilt
    int r = check_ready(c);
    if (r) return r;
    if (n && (!ids || !factors)) return c->fail(DEME_ERR_INVALID_ARGUMENT, "ids and factors must not be NULL");
    // per-clump factor (1 = unchanged); a repeated id takes its last factor
    std::vector<float> f(c->nClumps, 1.f);
    std::vector<uint8_t> mark(c->nClumps, 0);
    for (uint32_t k = 0; k < n; k++) {
        if (ids[k] >= c->nClumps) return c->fail(DEME_ERR_INVALID_ARGUMENT, "an id is not a clump owner");
        if (!std::isfinite(factors[k]) || !(factors[k] > 0.f))
            return c->fail(DEME_ERR_INVALID_ARGUMENT, "size factors must be finite and positive");
        f[ids[k]] = factors[k];
        mark[ids[k]] = 1;
    }
    if (!n) return DEME_OK;
    DBuf<float> df;
    DBuf<uint8_t> dm;
    DEME_HIP(c, df.ensure(c->nClumps));
    hipError_t e = dm.ensure(c->nClumps);
    if (e != hipSuccess) {
        df.release();
        return c->fail_hip(e, "change_clump_sizes");
    }
    e = hipMemcpyAsync(df.p, f.data(), c->nClumps * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dm.p, mark.data(), c->nClumps, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = launch_scale_components(c->sph.p, c->nS, dm.p, df.p, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    df.release();
    dm.release();
    if (e != hipSuccess) return c->fail_hip(e, "change_clump_sizes");
    // the host's copy of the radii, scaled by the same float products, sizes the cell grid
    std::vector<uint32_t> owner(c->nS);
    if (c->nS) e = hipMemcpy2D(owner.data(), 4, c->sph.p, sizeof(SphereRec), 4, c->nS, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return c->fail_hip(e, "change_clump_sizes");
    float maxR = 0.f;
    for (uint32_t i = 0; i < c->nS; i++) {
        if (mark[owner[i]]) c->hostRadius[i] *= f[owner[i]];
        maxR = std::max(maxR, c->hostRadius[i]);
    }
    if (maxR != c->maxRadius) c->nCells = 0;
    c->maxRadius = maxR;
    c->meanRadius = mean_radius(c->hostRadius);
    c->adjReady = false;
    c->graphGen++;
    return DEME_OK;
}

// ---------------------------------------------------------------------------------------------------------------
// Inspectors
// ---------------------------------------------------------------------------------------------------------------
int deme_set_owner_volumes(deme_ctx* c, const float* volume) {
    if (!c->haveTemplates) return c->fail(DEME_ERR_NOT_READY, "templates not uploaded (deme_upload_templates)");
    if (!volume && c->nClumps) return c->fail(DEME_ERR_INVALID_ARGUMENT, "volume is NULL");
    DEME_HIP(c, c->ownerVol.ensure(std::max<uint32_t>(c->nClumps, 1)));
    if (c->nClumps)
        DEME_HIP(c, hipMemcpyAsync(c->ownerVol.p, volume, (size_t)c->nClumps * sizeof(float), hipMemcpyHostToDevice,
                                   c->stream));
    DEME_HIP(c, hipStreamSynchronize(c->stream));
    c->haveVolumes = true;
    return DEME_OK;
}

int deme_inspect(deme_ctx* c, const char* quantity, const char* region, float* value) {
    int r = check_ready(c);
    if (r) return r;
    if (!quantity || !value) return c->fail(DEME_ERR_INVALID_ARGUMENT, "quantity and value must not be NULL");
    // DEMInspector::switch_quantity_type (AuxClasses.cpp:58-103)
    const std::string q = quantity;
    int what, flavor;
    if (q == "clump_max_z") {
        what = INSPECT_CLUMP_MAX_Z;
        flavor = INSPECT_MAX;
    } else if (q == "clump_min_z") {
        what = INSPECT_CLUMP_MIN_Z;
        flavor = INSPECT_MIN;
    } else if (q == "clump_max_absv") {
        what = INSPECT_CLUMP_MAX_ABSV;
        flavor = INSPECT_MAX;
    } else if (q == "clump_volume") {
        what = INSPECT_CLUMP_VOLUME;
        flavor = INSPECT_SUM;
        if (!c->haveVolumes) return c->fail(DEME_ERR_NOT_READY, "clump volumes are not uploaded");
    } else {
        return c->fail(DEME_ERR_INVALID_ARGUMENT, q + " is not a known query type.");
    }
    const std::string reg = region ? region : "";
    const bool all = reg.find_first_not_of(" \t\r\n") == std::string::npos;
    if (!all && (!(deme_jitgen::whole_word(reg, "X") || deme_jitgen::whole_word(reg, "Y") ||
                   deme_jitgen::whole_word(reg, "Z")) ||
                 !deme_jitgen::whole_word(reg, "return")))
        // the reference's check of the region code (AuxClasses.cpp:121-131)
        return c->fail(DEME_ERR_INVALID_ARGUMENT,
                       "the region of an inspector needs to return a bool that is a result of logical operations "
                       "involving X, Y and Z");
    const uint32_t n = what == INSPECT_CLUMP_VOLUME ? c->nClumps : c->nS;
    const uint32_t nb = inspect_blocks(n);
    DEME_HIP(c, c->inspPart.ensure((size_t)nb + 1));
    if (what != INSPECT_CLUMP_VOLUME) DEME_HIP(c, c->ownerVol.ensure(1));
    if (all) {
        DEME_LAUNCH(c, "inspect",
                    launch_inspect(c->kp, what, flavor, n, c->sph.p, c->own.p, c->ownerVol.p, c->sp.LBF,
                                   c->inspPart.p, c->stream));
    } else {
        auto it = c->inspectJit.find(reg);
        if (it == c->inspectJit.end()) {
            hipDeviceProp_t prop;
            DEME_HIP(c, hipGetDeviceProperties(&prop, c->device));
            std::vector<char> code;
            std::string log;
            if (jit_compile(jit_inspect_source(reg), prop.gcnArchName, code, log))
                return c->fail(DEME_ERR_JIT, "hipRTC compilation of the inspector region failed:\n" + log);
            JitInspect k;
            DEME_HIP(c, k.load(code));
            it = c->inspectJit.emplace(reg, k).first;
        }
        const JitInspect& k = it->second;
        DEME_LAUNCH(c, "inspect",
                    k.launch(c->kp, what, flavor, n, c->sph.p, c->own.p, c->ownerVol.p, c->sp.LBF, c->inspPart.p,
                             c->stream));
    }
    DEME_LAUNCH(c, "inspect_final", launch_inspect_final(flavor, nb, c->inspPart.p, c->inspPart.p + nb, c->stream));
    double out = 0.0;
    DEME_HIP(c, hipMemcpyAsync(&out, c->inspPart.p + nb, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    DEME_HIP(c, hipStreamSynchronize(c->stream));
    *value = (float)out;
    return DEME_OK;
}

int deme_set_time(deme_ctx* c, double t) {
    c->time = t;
    return DEME_OK;
}

int deme_enable_kernel_timing(deme_ctx* c, int on) {
    if (!on) {
        hipError_t e = drain_timing(c);
        if (e != hipSuccess) return c->fail_hip(e, "timing");
    }
    c->timing = on != 0;
    return DEME_OK;
}

int deme_reset_kernel_timing(deme_ctx* c) {
    hipError_t e = drain_timing(c);
    if (e != hipSuccess) return c->fail_hip(e, "timing");
    std::fill(c->ktotal.begin(), c->ktotal.end(), 0.0);
    std::fill(c->kcount.begin(), c->kcount.end(), 0);
    return DEME_OK;
}

int deme_kernel_timing(deme_ctx* c, int max, const char** names, double* total_ms, uint64_t* launches, int* n) {
    hipError_t e = drain_timing(c);
    if (e != hipSuccess) return c->fail_hip(e, "timing");
    const int k = (int)c->knames.size();
    for (int i = 0; i < k && i < max; i++) {
        if (names) names[i] = c->knames[i].c_str();
        if (total_ms) total_ms[i] = c->ktotal[i];
        if (launches) launches[i] = c->kcount[i];
    }
    *n = k;
    return DEME_OK;
}

}  // extern "C"
