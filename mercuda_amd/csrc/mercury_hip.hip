// libmercury_hip: host orchestration and the C ABI (include/mercury_hip.h).
//
// One global context per process, as the reference's CUDA backend
// (mercury_cuda.cu:16-32).  Device layout: component-major SoA over all
// bodies, body 0 the central body, Big bodies 1..nbig-1, particles after.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mercury_hip.h"
#include "mvs_device.hpp"
#include "radau_device.hpp"
#include "tangent.hpp"
#include "hybrid_device.hpp"
#include "bs_device.hpp"

using namespace mhip;

namespace {

void check(hipError_t code, const char* what) {
    if (code != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(code));
}
#define HIPCHECK(call) check((call), #call)

int fail(const std::exception& e) {
    std::fprintf(stderr, "MERCURY HIP: %s\n", e.what());
    return 1;
}

template <class T>
void dalloc(T*& p, size_t count) {
    if (p) HIPCHECK(hipFree(p));
    p = nullptr;
    if (count) HIPCHECK(hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * count));
}
template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

struct Timing {
    bool on = false;
    bool fast_only = false;   // time the fast particle kernel's launches only
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> marks;   // kind, (start, stop)
    size_t used = 0;
    hipEvent_t get() {
        if (used == pool.size()) {
            hipEvent_t e;
            HIPCHECK(hipEventCreate(&e));
            pool.push_back(e);
        }
        return pool[used++];
    }
    void clear() { marks.clear(); used = 0; }
};

constexpr int kRing = 8;   // launches in flight: table, counter and request slots
// control words of an MVS call (Context::dcnt)
constexpr int kCtlCount = 0;               // entries of slow-set list 0 and list 1
constexpr int kCtlPar = 2;                 // index of the current list
constexpr int kCtlLeave = 3;               // bodies leaving the slow set after the call
constexpr int kCtlPers = 4;                // kept members at the call's start
constexpr int kCtlSnap = 5;                // [kRing] current-list entries after launch k's fast kernel
constexpr int kCtlReq = 5 + kRing;         // [kRing] screening requests of ring slot k
constexpr int kCtlInts = 5 + 2 * kRing;

struct Context {
    int algorithm = 1;
    int n = 0, nbig = 0, pn = 0, ngflag = 0;
    double jcen[3] = {0, 0, 0};
    long long cap = 0;
    bool history = false;            // accelerations valid (dtflag == 2)
    bool saved_valid = false;        // step-start copies hold the last step
    bool have_state = false;
    // device state
    double *x = nullptr, *v = nullptr, *a = nullptr, *m = nullptr, *rce = nullptr, *rphys = nullptr;
    double *ox = nullptr, *ov = nullptr, *v0k = nullptr, *v1 = nullptr;
    double *cx = nullptr, *cv = nullptr;   // corrector output (allocated on first export)
    int* fault = nullptr;
    double* tab = nullptr;
    long long tab_cap = 0;           // doubles
    Sys* dsys = nullptr;
    int2* dtiles = nullptr;
    int ntiles = 0;
    DevEvent* dev = nullptr;
    int* devcount = nullptr;
    int ev_cap = 0;
    Park* dpark = nullptr;           // the two slow-set lists (park_cap entries each)
    int* dleave = nullptr;           // bodies leaving the slow set
    int* dcnt = nullptr;             // control words (kCtl*)
    unsigned char* dslow = nullptr;  // slow-set flag per body
    int park_cap = 0;
    // non-gravitational model (ngflag != 0): parameters [5][ngf_cap] (A1, A2, A3, beta, yar)
    double* dngf = nullptr;
    long long ngf_cap = 0;
    // MVS with the cometary model: the particles with cometary or Yarkovsky
    // parameters (global indices), permanent slow-set members, and their
    // accelerations angf [3][angf_cap] carried between steps
    std::vector<int> comets;
    Park* dseed = nullptr;
    int seed_cap = 0;
    double* dangf = nullptr;
    long long angf_cap = 0;
    int* hstatus = nullptr;          // pinned host words the call's last kernel writes (call_end)
    int* dstatus = nullptr;          // their device address
    unsigned upload_gen = 0;         // bumped by every upload (captured graphs depend on it)
    ScreenReq* dreq = nullptr;       // kRing slots of req_cap requests
    int req_cap = 0;
    long long slot_doubles = 0;      // table doubles per ring slot
    hipStream_t sp = nullptr, sg = nullptr;   // planet and general-kernel streams
    hipStream_t sm = nullptr;        // the MVS pipeline's particle stream (the main stream)
    hipEvent_t ev_plan[kRing + 1] = {}, ev_fast[kRing + 1] = {}, ev_done[kRing + 1] = {};
    hipEvent_t ev_gate = nullptr;    // each fast launch waits for it (recorded after a marker)
    // host
    std::vector<double> stage;
    std::vector<DevEvent> hev;
    std::vector<MercuryEvent> events;
    std::vector<MercuryHipEvent> adv_events;
    struct HostSys { int base, n, nbig; double mu, h, rcen, j2, j4, j6; };
    std::vector<HostSys> systems;    // one entry, or the members of an ensemble
    bool ensemble = false;           // systems came from ensemble_upload (each carries its h, rcen)
    std::vector<Sys> hsys;           // device descriptors last uploaded
    bool sys_valid = false;
    int sys_cap = 0;
    double* tcur = nullptr;          // device per-system time
    int npart = 0;                   // particles over all systems
    hipStream_t stream = nullptr;
    int steps_per_launch = 32;
    int block = 128;
    int ppl = 1;          // fast kernel: particles per lane
    int gen_per_wave = 4;
    int64_t parked = 0;
    unsigned long long* work = nullptr;   // device particle-step counters (timing mode)
    long long work_total = 0;             // particle-steps of the MVS calls since timing was enabled
    Timing timing;
    // RA15 state (algorithm 4): b, e [21][cap], a1 [3][cap], max |b7|
    double *rb = nullptr, *re = nullptr, *ra1 = nullptr;
    // the last sequence's finish, not yet applied (the next sequence's kernels
    // apply it, or radau_flush when the state is read first)
    bool ra_pend = false, ra_preset = false;
    double ra_ptdid = 0.0, ra_pq = 0.0;
    // whole-system Bulirsch-Stoer (algorithms 2, 3): accepted states per level, level errors
    double* bscand = nullptr;
    double* bsrsave = nullptr;
    long long bs_cap = 0;
    unsigned long long* bserr = nullptr;
    unsigned long long* bmax = nullptr;
    long long ra_cap = 0;
    // HYBRID (algorithm 10): changeover radii, pre-drift states, tables, pairs
    double *rcrit = nullptr, *hx0 = nullptr, *hv0 = nullptr, *hytab = nullptr, *hpack = nullptr;
    int *hpairs = nullptr, *hnpairs = nullptr, *hids = nullptr;
    int hpair_cap = 0, hpack_cap = 0;
    long long hy_cap = 0;
    std::vector<int> members;          // encounter members of the current step (body indices)
    std::vector<double> packed;
    // compact encounter context (mdt_hkce's BS2 set)
    struct Enc {
        bool active = false;
        int n = 0, nbig = 0, ncap = 0;
        double *m = nullptr, *rcrit = nullptr, *x = nullptr, *v = nullptr, *tp = nullptr, *cand = nullptr;
        int *bbp = nullptr, *poff = nullptr, *pidx = nullptr;
        int4* pinfo = nullptr;         // packed partners per particle (pack_partners)
        double4* pcon = nullptr;
        int nbbp = 0, npidx_cap = 0, nbbp_cap = 0;
        unsigned long long* err = nullptr;
        BsCtl* ctl = nullptr;          // device step control (encounter_advance)
        std::vector<double> stage;
        // pinned host words of a per-substep attempt (12 level errors, the
        // fault flags) and the pinned staging of a state download, with the
        // device addresses the kernels write them through
        unsigned long long* hbits = nullptr;
        unsigned long long* dbits = nullptr;
        double* hstage = nullptr;
        double* dstage = nullptr;
        size_t hstage_cap = 0;
        bool pending = false;          // an accepted substep's kernel may still run on the stream
    } enc;
    // differentiation tape (mercury_hip_adjoint_forward)
    struct Tape {
        bool valid = false;
        int steps = 0, seg = 0, nseg = 0, npart = 0, np = 0, nblk = 0;
        Sys hs{};
        Sys* dsys = nullptr;
        double *tab = nullptr, *ck = nullptr, *rec = nullptr, *lam = nullptr, *part = nullptr, *src = nullptr;
        double *grad = nullptr, *rows = nullptr, *tan = nullptr, *fin = nullptr, *seed = nullptr;
        double *psave = nullptr;
        double *pfin = nullptr;           // final planet tangents per direction [7np][6np]
        std::vector<double> xfin, vfin;   // final states (host, Fortran layout)
    } tape;
};

Context ctx;

// ---- allocation-free force (mercury_hip_force_noalloc) ---------------------
// Its kernels' scratch lives in this file-scope array, one copy per device,
// held until the process ends: the system descriptor at byte 0, the fault flag
// at byte 256, the planets' force table from byte 512 (kNoAllocTable doubles:
// the header and force block of up to kMaxBig - 1 planets).  Calls on one
// device share it: calls that overlap on one device race.
constexpr size_t kNoAllocBytes = 4096;
constexpr size_t kNoAllocFault = 256, kNoAllocTab = 512;
constexpr int kNoAllocTable = (int)((kNoAllocBytes - kNoAllocTab) / sizeof(double));
static_assert(sizeof(Sys) <= kNoAllocFault, "descriptor slot");
static_assert(kHdr + kF * (kMaxBig - 1) <= kNoAllocTable, "table slot");
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];
constexpr int kMaxDevices = 64;
unsigned char* scratch_ptr[kMaxDevices] = {};   // device address of `scratch` on each device

// The descriptor comes by value (no host-to-device copy); the fault flag and
// the central body's output row are cleared.
__global__ void noalloc_setup(Sys d, unsigned char* scr, double* a, long long ld) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *reinterpret_cast<Sys*>(scr) = d;
        *reinterpret_cast<int*>(scr + kNoAllocFault) = 0;
        a[0] = 0.0; a[ld] = 0.0; a[2 * ld] = 0.0;
    }
}

// What a captured advance depends on: replay only when all of it is unchanged.
struct GraphKey {
    int nsteps, screen, reset, spl, block, ppl, gpw, ntiles, nsys, ev_cap, req_cap, park_cap, n, npart;
    unsigned gen;
    long long cap, slot;
    const void* ptr[11];
    bool operator==(const GraphKey& o) const { return std::memcmp(this, &o, sizeof(GraphKey)) == 0; }
};

struct GraphCache {
    bool enabled = false;
    bool valid = false;
    GraphKey key{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int captures = 0, replays = 0;
    void drop() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        valid = false;
    }
} graphs;

void ensure_stream() {
    if (ctx.stream) return;
    if (const char* e = std::getenv("MERCURY_HIP_GEN_PER_WAVE")) ctx.gen_per_wave = std::max(1, std::min(64, std::atoi(e)));
    HIPCHECK(hipStreamCreateWithFlags(&ctx.stream, hipStreamNonBlocking));
    ctx.sm = ctx.stream;
    HIPCHECK(hipStreamCreateWithFlags(&ctx.sp, hipStreamNonBlocking));
    HIPCHECK(hipStreamCreateWithFlags(&ctx.sg, hipStreamNonBlocking));
    if (!ctx.hstatus) {
        HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&ctx.hstatus), 64, hipHostMallocDefault));
        std::memset(ctx.hstatus, 0, 64);
        HIPCHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&ctx.dstatus), ctx.hstatus, 0));
    }
    HIPCHECK(hipEventCreateWithFlags(&ctx.ev_gate, hipEventDisableTiming));
    for (int k = 0; k <= kRing; k++) {
        HIPCHECK(hipEventCreateWithFlags(&ctx.ev_plan[k], hipEventDisableTiming));
        HIPCHECK(hipEventCreateWithFlags(&ctx.ev_fast[k], hipEventDisableTiming));
        HIPCHECK(hipEventCreateWithFlags(&ctx.ev_done[k], hipEventDisableTiming));
    }
}

void sync_all() {
    HIPCHECK(hipStreamSynchronize(ctx.sm));
    HIPCHECK(hipStreamSynchronize(ctx.sp));
    HIPCHECK(hipStreamSynchronize(ctx.stream));
    HIPCHECK(hipStreamSynchronize(ctx.sg));
}

void release_all() {
    Context& c = ctx;
    if (c.sg) (void)hipStreamSynchronize(c.sg);   // a tape's tangent pass may still run
    if (c.stream) (void)hipStreamSynchronize(c.stream);   // an encounter substep's accept may still run
    c.enc.pending = false;
    if (c.enc.hbits) (void)hipHostFree(c.enc.hbits);
    if (c.enc.hstage) (void)hipHostFree(c.enc.hstage);
    c.enc.hbits = nullptr; c.enc.hstage = nullptr; c.enc.hstage_cap = 0;
    c.enc.dbits = nullptr; c.enc.dstage = nullptr;
    graphs.drop();
    for (double** p : {&c.x, &c.v, &c.a, &c.m, &c.rce, &c.rphys, &c.ox, &c.ov, &c.v0k, &c.v1, &c.tab, &c.cx, &c.cv})
        dfree(*p);
    dfree(c.fault); dfree(c.dsys); dfree(c.tcur); c.sys_valid = false; c.sys_cap = 0; c.systems.clear(); dfree(c.dtiles); dfree(c.dev); dfree(c.devcount);
    dfree(c.dpark); dfree(c.dleave); dfree(c.dcnt); dfree(c.dslow); dfree(c.work); dfree(c.dreq); c.req_cap = 0;
    dfree(c.rb); dfree(c.re); dfree(c.ra1); dfree(c.bmax); c.ra_cap = 0; c.ra_pend = false;
    dfree(c.bscand); dfree(c.bsrsave); dfree(c.bserr); c.bs_cap = 0;
    dfree(c.dngf); c.ngf_cap = 0; dfree(c.dangf); c.angf_cap = 0; dfree(c.dseed); c.seed_cap = 0;
    c.comets.clear(); c.ngflag = 0;
    for (double** p : {&c.tape.tab, &c.tape.ck, &c.tape.rec, &c.tape.lam, &c.tape.part, &c.tape.src, &c.tape.grad,
                       &c.tape.rows, &c.tape.tan, &c.tape.fin, &c.tape.seed, &c.tape.psave, &c.tape.pfin})
        dfree(*p);
    dfree(c.tape.dsys);
    c.tape.valid = false;
    for (double** p : {&c.rcrit, &c.hx0, &c.hv0, &c.hytab, &c.hpack, &c.enc.m, &c.enc.rcrit, &c.enc.x, &c.enc.v,
                       &c.enc.tp, &c.enc.cand})
        dfree(*p);
    dfree(c.hpairs); dfree(c.hnpairs); dfree(c.hids); dfree(c.enc.bbp); dfree(c.enc.poff); dfree(c.enc.pidx);
    dfree(c.enc.err); dfree(c.enc.ctl); dfree(c.enc.pinfo); dfree(c.enc.pcon);
    c.hpair_cap = 0; c.hpack_cap = 0; c.hy_cap = 0; c.enc.ncap = 0; c.enc.active = false;
    c.enc.npidx_cap = 0; c.enc.nbbp_cap = 0;
    c.cap = 0; c.tab_cap = 0; c.ev_cap = 0; c.park_cap = 0; c.ntiles = 0;
    c.n = 0; c.nbig = 0; c.history = false; c.saved_valid = false; c.have_state = false;
    c.events.clear(); c.adv_events.clear();
}

State device_state() {
    State s;
    s.x = ctx.x; s.v = ctx.v; s.a = ctx.a; s.m = ctx.m; s.rce = ctx.rce; s.rphys = ctx.rphys;
    s.ox = ctx.ox; s.ov = ctx.ov; s.v0k = ctx.v0k; s.v1 = ctx.v1; s.fault = ctx.fault; s.cap = ctx.cap;
    s.ngf = (ctx.ngflag && !ctx.ensemble) ? ctx.dngf : nullptr;
    s.angf = ctx.dangf;
    return s;
}

void ensure_capacity(int n) {
    if (n <= ctx.cap) return;
    long long cap = ((long long)n + 255) / 256 * 256;
    size_t free_b = 0, total_b = 0;
    HIPCHECK(hipMemGetInfo(&free_b, &total_b));
    if ((size_t)cap * 8 * 40 > free_b) throw std::runtime_error("insufficient device memory");
    for (double** p : {&ctx.x, &ctx.v, &ctx.a, &ctx.ox, &ctx.ov, &ctx.v0k, &ctx.v1}) dalloc(*p, 3 * cap);
    dfree(ctx.cx);
    dfree(ctx.cv);
    for (double** p : {&ctx.m, &ctx.rce, &ctx.rphys}) dalloc(*p, cap);
    if (!ctx.fault) dalloc(ctx.fault, 1);
    ctx.park_cap = (int)std::min<long long>(cap, 1LL << 29);
    dalloc(ctx.dpark, (size_t)2 * ctx.park_cap);
    dalloc(ctx.dleave, (size_t)ctx.park_cap);
    if (!ctx.dcnt) dalloc(ctx.dcnt, kCtlInts);
    dalloc(ctx.dslow, (size_t)cap);
    ctx.req_cap = (int)std::min<long long>(std::max<long long>(cap / 2, 1 << 18), 1LL << 26);
    dalloc(ctx.dreq, (size_t)ctx.req_cap * kRing);
    if (!ctx.devcount) dalloc(ctx.devcount, 1);
    if (!ctx.dev) {
        // the event list starts at 2^18 records (MERCURY_HIP_EVENT_CAP overrides)
        // and doubles when a step's events need more (mercury_hip_events)
        ctx.ev_cap = 1 << 18;
        if (const char* e = std::getenv("MERCURY_HIP_EVENT_CAP")) ctx.ev_cap = std::max(1, std::atoi(e));
        dalloc(ctx.dev, (size_t)ctx.ev_cap);
    }
    ctx.cap = cap;
}

__global__ void seed_slow(const Park* seeds, int n, unsigned char* slow) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) slow[seeds[i].j] = 5;
}

// Empties the slow set (a new state, or one put back): every particle starts
// the next call in the fast kernel, except the comets of the cometary model,
// which are permanent members: they enter the list the next call's
// call_begin makes current (list 1 after the cleared control words' list 0)
// as kept members.
void clear_slow_set() {
    if (ctx.dslow) HIPCHECK(hipMemset(ctx.dslow, 0, (size_t)ctx.cap));
    if (ctx.dcnt) HIPCHECK(hipMemset(ctx.dcnt, 0, sizeof(int) * kCtlInts));
    const int nc = (int)ctx.comets.size();
    if (nc == 0 || ctx.ensemble || !ctx.dpark) return;
    if (nc > ctx.park_cap) throw std::runtime_error("more comets than the slow-set capacity");
    std::vector<Park> seeds(nc);
    for (int i = 0; i < nc; i++) seeds[i] = Park{ctx.comets[i], 0, 0, -1};
    if (ctx.seed_cap < nc) {
        dalloc(ctx.dseed, (size_t)nc);
        ctx.seed_cap = nc;
    }
    HIPCHECK(hipMemcpy(ctx.dseed, seeds.data(), sizeof(Park) * nc, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(ctx.dpark + ctx.park_cap, ctx.dseed, sizeof(Park) * nc, hipMemcpyDeviceToDevice));
    HIPCHECK(hipMemcpy(ctx.dcnt + kCtlCount + 1, &nc, sizeof(int), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(seed_slow, dim3(std::min((nc + 255) / 256, 1024)), dim3(256), 0, 0, ctx.dseed, nc, ctx.dslow);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipDeviceSynchronize());
}

void ensure_table(long long doubles) {
    if (doubles <= ctx.tab_cap) return;
    dalloc(ctx.tab, (size_t)doubles);
    ctx.tab_cap = doubles;
}

void ensure_events(int count) {
    if (count <= ctx.ev_cap) return;
    int cap = ctx.ev_cap;
    while (cap < count) cap *= 2;
    dalloc(ctx.dev, (size_t)cap);
    ctx.ev_cap = cap;
}

void ensure_systems(int nsys) {
    if (nsys <= ctx.sys_cap) return;
    dalloc(ctx.dsys, (size_t)nsys);
    dalloc(ctx.tcur, (size_t)nsys);
    ctx.sys_cap = nsys;
    ctx.sys_valid = false;
}

// Workgroup tiles over the particles of every system, block * ppl particles
// each; a tile never spans two systems, so each workgroup reads one system's
// descriptor and table.
void build_tiles() {
    int b = ctx.block;
    std::vector<int2> t;
    int npart = 0;
    for (size_t s = 0; s < ctx.systems.size(); s++) {
        const auto& hs = ctx.systems[s];
        for (int j = hs.base + hs.nbig; j < hs.base + hs.n; j += b * ctx.ppl) t.push_back(make_int2((int)s, j));
        npart += hs.n - hs.nbig;
    }
    dalloc(ctx.dtiles, std::max<size_t>(t.size(), 1));
    if (!t.empty())
        HIPCHECK(hipMemcpy(ctx.dtiles, t.data(), sizeof(int2) * t.size(), hipMemcpyHostToDevice));
    ctx.ntiles = (int)t.size();
    ctx.npart = npart;
}

// Fortran x(3,n) -> device SoA [3][cap]
void upload_vec3(const double* src, double* dst) {
    int n = ctx.n;
    ctx.stage.resize((size_t)3 * n);
    for (int j = 0; j < n; j++)
        for (int k = 0; k < 3; k++) ctx.stage[(size_t)k * n + j] = src[3 * j + k];
    for (int k = 0; k < 3; k++)
        HIPCHECK(hipMemcpy(dst + k * ctx.cap, ctx.stage.data() + (size_t)k * n, sizeof(double) * n,
                           hipMemcpyHostToDevice));
}

void download_vec3(const double* src, double* dst) {
    int n = ctx.n;
    ctx.stage.resize((size_t)3 * n);
    for (int k = 0; k < 3; k++)
        HIPCHECK(hipMemcpy(ctx.stage.data() + (size_t)k * n, src + k * ctx.cap, sizeof(double) * n,
                           hipMemcpyDeviceToHost));
    for (int j = 0; j < n; j++)
        for (int k = 0; k < 3; k++) dst[3 * j + k] = ctx.stage[(size_t)k * n + j];
}

void timed_launch_begin(hipStream_t st, hipEvent_t& start, int kind = 0) {
    if (!ctx.timing.on || (ctx.timing.fast_only && kind != 0)) return;
    start = ctx.timing.get();
    HIPCHECK(hipEventRecord(start, st));
}
void timed_launch_end(int kind, hipStream_t st, hipEvent_t start) {
    if (!ctx.timing.on || (ctx.timing.fast_only && kind != 0)) return;
    hipEvent_t stop = ctx.timing.get();
    HIPCHECK(hipEventRecord(stop, st));
    ctx.timing.marks.push_back({kind, {start, stop}});
}

enum { K_PARTICLE = 0, K_PLANET = 1, K_GENERAL = 2 };

// Device descriptors of every system for launches of at most smax steps.
// A single uploaded system takes h and rcen from the call (h_call > 0 flag);
// ensemble members carry their own.  Table block of one launch: every
// system's step tables, then every system's start table.
void build_systems(int smax, bool use_call, double h_call, double rcen_call) {
    Context& c = ctx;
    const int nsys = (int)c.systems.size();
    std::vector<Sys> ds(nsys);
    long long off = 0;
    for (int s = 0; s < nsys; s++) {
        const auto& hs = c.systems[s];
        int np = std::max(hs.nbig - 1, 0);
        Sys d{};
        d.base = hs.base; d.n = hs.n; d.nbig = hs.nbig; d.ts = table_stride(np);
        d.tab = off;
        off += (long long)d.ts * smax;
        d.mu = hs.mu;
        d.h = use_call ? h_call : hs.h;
        d.hby2 = .5 * d.h;
        d.rcen = use_call ? rcen_call : hs.rcen;
        d.j2 = hs.j2; d.j4 = hs.j4; d.j6 = hs.j6;
        d.obl = (hs.j2 != 0 || hs.j4 != 0 || hs.j6 != 0);
        d.ngflag = c.ensemble ? 0 : c.ngflag;
        ds[s] = d;
    }
    for (int s = 0; s < nsys; s++) {
        ds[s].start = off;
        off += start_stride(std::max(c.systems[s].nbig - 1, 0));
    }
    c.slot_doubles = (off + 63) / 64 * 64;
    ensure_table(c.slot_doubles * kRing);
    ensure_systems(nsys);
    if (c.sys_valid && c.hsys.size() == ds.size() &&
        std::memcmp(c.hsys.data(), ds.data(), sizeof(Sys) * ds.size()) == 0)
        return;
    c.hsys = ds;
    HIPCHECK(hipMemcpy(c.dsys, c.hsys.data(), sizeof(Sys) * ds.size(), hipMemcpyHostToDevice));
    c.sys_valid = true;
}

// The fast kernel specialises on each workgroup's planet count itself (see
// mvs_particles_fast); np is not needed here.
void launch_fast(bool screen, bool save, int, dim3 grid, dim3 block, hipStream_t st, const ParticleArgs& args) {
    // waves per SIMD the fast kernel's registers are capped for: 5 measured
    // best on MI355X (a few spills, more latency hidden); 1 leaves the
    // allocation to the compiler
    static const int occ = std::getenv("MERCURY_HIP_FAST_OCC") ? std::atoi(std::getenv("MERCURY_HIP_FAST_OCC")) : 5;
    const size_t shm = screen ? sizeof(double) * kEnds * block.x : 0;
    if (occ == 1) {
        if (screen && save) hipLaunchKernelGGL((mvs_particles_fast<true, true, 1>), grid, block, shm, st, args);
        else if (screen) hipLaunchKernelGGL((mvs_particles_fast<true, false, 1>), grid, block, shm, st, args);
        else if (save) hipLaunchKernelGGL((mvs_particles_fast<false, true, 1>), grid, block, shm, st, args);
        else hipLaunchKernelGGL((mvs_particles_fast<false, false, 1>), grid, block, shm, st, args);
    } else {
        if (screen && save) hipLaunchKernelGGL((mvs_particles_fast<true, true, 5>), grid, block, shm, st, args);
        else if (screen) hipLaunchKernelGGL((mvs_particles_fast<true, false, 5>), grid, block, shm, st, args);
        else if (save) hipLaunchKernelGGL((mvs_particles_fast<false, true, 5>), grid, block, shm, st, args);
        else hipLaunchKernelGGL((mvs_particles_fast<false, false, 5>), grid, block, shm, st, args);
    }
    HIPCHECK(hipGetLastError());
}
void launch_general(bool screen, bool save, bool uniform, dim3 grid, dim3 block, hipStream_t st,
                    const ParticleArgs& args) {
    if (uniform && !ctx.comets.empty()) {   // one system with the cometary model's members
        if (screen && save) hipLaunchKernelGGL((mvs_particles_general<true, true, true, true>), grid, block, 0, st, args);
        else if (screen) hipLaunchKernelGGL((mvs_particles_general<true, false, true, true>), grid, block, 0, st, args);
        else if (save) hipLaunchKernelGGL((mvs_particles_general<false, true, true, true>), grid, block, 0, st, args);
        else hipLaunchKernelGGL((mvs_particles_general<false, false, true, true>), grid, block, 0, st, args);
    } else if (uniform) {   // one system: wave-uniform table reads
        if (screen && save) hipLaunchKernelGGL((mvs_particles_general<true, true, true>), grid, block, 0, st, args);
        else if (screen) hipLaunchKernelGGL((mvs_particles_general<true, false, true>), grid, block, 0, st, args);
        else if (save) hipLaunchKernelGGL((mvs_particles_general<false, true, true>), grid, block, 0, st, args);
        else hipLaunchKernelGGL((mvs_particles_general<false, false, true>), grid, block, 0, st, args);
    } else {
        if (screen) hipLaunchKernelGGL((mvs_particles_general<true, false, false>), grid, block, 0, st, args);
        else hipLaunchKernelGGL((mvs_particles_general<false, false, false>), grid, block, 0, st, args);
    }
    HIPCHECK(hipGetLastError());
}

// Launch lengths for nsteps with at most smax per launch.  Launch boundaries
// do not change results.
// Launch lengths of one call: a short first launch, so that the particles
// start after a few planet steps while the planet wave tabulates the next
// launch's steps, then launches growing by a factor (4, 16, 32, ... with the
// defaults).  The planet wave takes ~9 us per step and the particle launch
// ~30 us per step at 1e6 particles, so a launch of k steps covers the next
// one's ~3k planet steps.  MERCURY_HIP_PLAN_HEAD and MERCURY_HIP_PLAN_GROW
// override the two numbers (head 0: every launch smax steps).
std::vector<int> chunk_plan(int nsteps, int smax) {
    static const int head = std::getenv("MERCURY_HIP_PLAN_HEAD") ? std::atoi(std::getenv("MERCURY_HIP_PLAN_HEAD")) : 4;
    static const int grow = std::getenv("MERCURY_HIP_PLAN_GROW") ? std::atoi(std::getenv("MERCURY_HIP_PLAN_GROW")) : 4;
    std::vector<int> plan;
    int c = head > 0 ? std::min(head, smax) : smax;
    for (int rest = nsteps; rest > 0;) {
        const int k = std::min(c, rest);
        plan.push_back(k);
        rest -= k;
        c = std::min(smax, c * std::max(grow, 1));
    }
    return plan;
}

// nsteps MVS steps of the single system or ensemble in launches of at most
// `smax` steps.  First, on the planet stream, call_begin (slow-set rotation,
// counters).  Per launch k (ring slot k % kRing):
//   planet stream:   planets
//   general stream:  the slow set over launch k's steps, once their planet
//                    tables exist and launch k-1's fast kernel has ended (it
//                    first finishes launch k-1's parks): beside launch k's fast
//                    kernel; then launch k-1's deferred screening details
//   main stream:     the fast particle kernel once the launch's planet tables
//                    exist, then the slow-set count snapshot (not after the
//                    call's last one: the finishing launch reads the count)
// A last general launch finishes the last fast kernel's parks, beside that
// kernel's screening details on the main stream.  The
// latency-bound chains (planets, slow set) thus run beside the fast kernel
// instead of before or after it; the planet stream runs up to kRing launches
// ahead, and its launches are enqueued two launches ahead of the others (a
// call's first kernels would otherwise wait on the host enqueueing them).
// Last, the streams join the main stream and call_end writes the host's
// status words.  Enqueue only (no host synchronisation): usable under
// stream capture.  device_time: every planet launch reads the systems' clocks
// from the device (set beforehand) instead of taking the call's time as an
// argument; this is the captured form, whose planet stream first joins the
// capture.
void mvs_enqueue(double time, double h, int nsteps, bool reset, bool screen, bool save, bool device_time) {
    Context& c = ctx;
    const int smax = save ? 1 : c.steps_per_launch;
    const int nsys = (int)c.systems.size();
    const bool single = nsys == 1;
    State st = device_state();
    EventSink es{c.dev, c.devcount, c.ev_cap, c.fault};
    if (device_time) {
        HIPCHECK(hipEventRecord(c.ev_fast[kRing], c.sm));
        HIPCHECK(hipStreamWaitEvent(c.sp, c.ev_fast[kRing], 0));
    }
    int* const count2 = c.dcnt + kCtlCount;
    int* const par = c.dcnt + kCtlPar;
    int* const pers = c.dcnt + kCtlPers;
    CallBeginArgs cb{c.dslow, c.dleave, count2, par, pers, c.dcnt + kCtlLeave, c.dcnt + kCtlSnap, 2 * kRing,
                     c.devcount, c.fault};
    // one system: the first planet launch does call_begin's work itself (its
    // one workgroup before the planets); an ensemble's planet workgroups run
    // concurrently, so call_begin precedes them as a launch of its own
    if (!single) {
        hipLaunchKernelGGL(call_begin, dim3(1), dim3(256), 0, c.sp, cb);
        HIPCHECK(hipGetLastError());
    }
    const int npart = c.npart;
    // The slow-set and screening kernels grid-stride over counts only the device
    // knows (a few hundred slow-set members and tens of thousands of requests per
    // launch at 1M particles); small grids keep their idle workgroups from
    // competing with the fast kernel's for dispatch.
    const int gblocks = std::max(1, std::min((npart + c.gen_per_wave - 1) / c.gen_per_wave, 512));
    const int rblocks = std::max(1, std::min((c.req_cap + 255) / 256, 256));
    double t = time;
    int done = 0;
    const std::vector<int> plan = chunk_plan(nsteps, smax);
    const int nk = (int)plan.size();
    const bool parts = c.ntiles > 0;
    auto slot_tab = [&](int k) { return c.tab + (long long)(k % kRing) * c.slot_doubles; };
    auto park_list = [&](int chunk, bool last) {
        const int* snap = c.dcnt + kCtlSnap + ((chunk + kRing - 1) % kRing);   // after fast kernel chunk-1
        return ParkList{c.dpark, count2, par, pers, snap, c.dleave, c.dcnt + kCtlLeave, c.dslow, c.park_cap,
                        chunk, c.gen_per_wave, last ? 1 : 0};
    };
    auto req_sink = [&](int k) {
        return ReqSink{c.dreq + (size_t)(k % kRing) * c.req_cap, c.dcnt + kCtlReq + (k % kRing), c.req_cap, c.fault};
    };
    auto general = [&](const ParticleArgs& a) {
        hipEvent_t g0 = nullptr;
        timed_launch_begin(c.sg, g0, K_GENERAL);
        launch_general(screen, save, single, dim3(gblocks), dim3(64), c.sg, a);
        timed_launch_end(K_GENERAL, c.sg, g0);
    };
    auto screen_details = [&](int k, int step0, hipStream_t stream) {
        if (!screen) return;
        RequestArgs ra{c.dsys, st, slot_tab(k), es, req_sink(k), step0};
        hipLaunchKernelGGL(screen_requests, dim3(rblocks), dim3(256), 0, stream, ra);
        HIPCHECK(hipGetLastError());
    };
    // steps done before each launch
    std::vector<int> done_at(nk + 1, 0);
    for (int k = 0; k < nk; k++) done_at[k + 1] = done_at[k] + plan[k];
    // Planet launch j on the planet stream.  They are enqueued up to
    // kPlanAhead launches ahead of the particle launches, so the planet wave
    // never waits on the host enqueueing the launches between them; a ring
    // slot's reuse waits on an event recorded kRing - kPlanAhead iterations
    // earlier.
    constexpr int kPlanAhead = 2;
    static_assert(kPlanAhead + 1 <= kRing, "a slot's release is recorded before its reuse is enqueued");
    auto enqueue_planets = [&](int j) {
        const int slot = j % kRing;
        if (j >= kRing) HIPCHECK(hipStreamWaitEvent(c.sp, c.ev_done[slot], 0));
        const bool rs0 = reset && j == 0;
        // one system follows the caller's clock; ensemble members carry theirs on the device
        const bool set_time = !device_time && ((single && !c.ensemble) || j == 0);
        PlanetArgs pa{c.dsys, st, slot_tab(j), c.tcur, es, t, set_time ? 1 : 0, plan[j], rs0 ? 1 : 0,
                      screen ? 1 : 0, save ? 1 : 0, c.dcnt + kCtlReq + slot, 1, done_at[j]};
        if (single && j == 0) { pa.begin = 1; pa.cb = cb; }
        hipEvent_t t0 = nullptr;
        timed_launch_begin(c.sp, t0, K_PLANET);
        hipLaunchKernelGGL(mvs_planets, dim3(nsys), dim3(64), 0, c.sp, pa);
        HIPCHECK(hipGetLastError());
        timed_launch_end(K_PLANET, c.sp, t0);
        for (int q = 0; q < plan[j]; q++) t = t + h;
        HIPCHECK(hipEventRecord(c.ev_plan[slot], c.sp));
    };
    for (int j = 0; j < std::min(nk, kPlanAhead); j++) enqueue_planets(j);
    int prev_done = 0;
    for (int k = 0; k < nk; k++) {
        const int steps = plan[k];
        const int slot = k % kRing;
        double* tab = slot_tab(k);
        const bool rs0 = reset && k == 0;
        const bool last = k + 1 == nk;
        int* snap = c.dcnt + kCtlSnap + slot;
        hipEvent_t t0 = nullptr;
        ParticleArgs args{c.dsys, c.dtiles, st, tab, es, park_list(k, false), req_sink(k),
                          c.timing.on ? c.work : nullptr, steps, rs0 ? 1 : 0, screen ? 1 : 0, save ? 1 : 0, done};
        args.ppl = c.ppl;
        if (parts) {
            // the slow set over this launch's steps, beside its fast kernel
            ParticleArgs ga = args;
            ga.park = park_list(k, last);
            if (k > 0) {
                ga.ptab = slot_tab(k - 1);
                ga.psteps = plan[k - 1];
                HIPCHECK(hipStreamWaitEvent(c.sg, c.ev_fast[(k - 1) % kRing], 0));
            }
            HIPCHECK(hipStreamWaitEvent(c.sg, c.ev_plan[slot], 0));
            // Each fast launch starts behind a marker on the general stream: the
            // slow set's waves, dispatched right after it, then take their CU
            // slots before the fast kernel's fill every SIMD's registers (a slow-
            // set launch queued behind a running fast launch would otherwise wait
            // for it to drain).
            hipLaunchKernelGGL(marker, dim3(1), dim3(64), 0, c.sg);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipEventRecord(c.ev_gate, c.sg));
            const hipEvent_t gate = c.ev_gate;
            general(ga);
            if (k > 0) {
                // launch k-1's screening details; its table slot is free after them
                screen_details(k - 1, prev_done, c.sg);
                HIPCHECK(hipEventRecord(c.ev_done[(k - 1) % kRing], c.sg));
            }
            HIPCHECK(hipStreamWaitEvent(c.sm, gate, 0));
            timed_launch_begin(c.sm, t0);
            launch_fast(screen, save, single ? c.systems[0].nbig - 1 : -1, dim3(c.ntiles), dim3(c.block), c.sm,
                        args);
            timed_launch_end(K_PARTICLE, c.sm, t0);
            // the next launch's slow set reads the count of this launch's
            // parks from a snapshot (the fast kernel after it appends more);
            // the call's finishing launch reads it directly
            if (!last) hipLaunchKernelGGL(snap_count, dim3(1), dim3(64), 0, c.sm, count2, par, snap);
            HIPCHECK(hipEventRecord(c.ev_fast[slot], c.sm));
        } else {
            HIPCHECK(hipStreamWaitEvent(c.sg, c.ev_plan[slot], 0));
            HIPCHECK(hipEventRecord(c.ev_done[slot], c.sg));
        }
        prev_done = done;
        done += steps;
        if (k + kPlanAhead < nk) enqueue_planets(k + kPlanAhead);
    }
    if (parts && nk > 0) {
        // the last fast kernel's parks finish their block (general stream, the
        // live count); beside them, its screening details (main stream, behind
        // the fast kernel that wrote the requests)
        ParticleArgs fa{c.dsys, c.dtiles, st, slot_tab(nk), es, park_list(nk, true), req_sink(nk - 1),
                        c.timing.on ? c.work : nullptr, 0, 0, screen ? 1 : 0, save ? 1 : 0, done};
        fa.park.live = 1;
        fa.ptab = slot_tab(nk - 1);
        fa.psteps = plan[nk - 1];
        HIPCHECK(hipStreamWaitEvent(c.sg, c.ev_fast[(nk - 1) % kRing], 0));
        general(fa);
        screen_details(nk - 1, prev_done, c.sm);
        HIPCHECK(hipEventRecord(c.ev_done[(nk - 1) % kRing], c.sm));
    }
    // the planet and general streams join the main stream; then the status words
    HIPCHECK(hipEventRecord(c.ev_plan[kRing], c.sp));
    HIPCHECK(hipEventRecord(c.ev_done[kRing], c.sg));
    HIPCHECK(hipStreamWaitEvent(c.sm, c.ev_plan[kRing], 0));
    HIPCHECK(hipStreamWaitEvent(c.sm, c.ev_done[kRing], 0));
    hipLaunchKernelGGL(call_end, dim3(1), dim3(64), 0, c.sm, count2, par, c.devcount, c.fault, c.dstatus);
    HIPCHECK(hipGetLastError());
}

__global__ void set_clocks(double* tcur, int nsys, double time) {
    for (int s = threadIdx.x; s < nsys; s += blockDim.x) tcur[s] = time;
}


void mvs_run(double time, double h, int nsteps, bool reset, bool screen, bool save, double rcen) {
    Context& c = ctx;
    const int smax = save ? 1 : c.steps_per_launch;
    const int nsys = (int)c.systems.size();
    const bool single = nsys == 1;
    if (save && !single) throw std::runtime_error("single-step mode needs one system");
    // an uploaded system follows the call's h and rcen; ensemble members carry theirs
    build_systems(smax, single && !c.ensemble, h, rcen);
    if (c.timing.on) c.work_total += (long long)c.npart * nsteps;
    if (!graphs.enabled || save || c.timing.on) {
        mvs_enqueue(time, h, nsteps, reset, screen, save, false);
        HIPCHECK(hipStreamSynchronize(c.sm));   // every stream has joined it
        return;
    }
    // hipGraph: capture the call's whole launch sequence once, replay it for
    // later calls of the same shape; the clocks come from the device
    GraphKey k;
    std::memset(&k, 0, sizeof(k));
    k.nsteps = nsteps; k.screen = screen; k.reset = reset; k.spl = c.steps_per_launch; k.block = c.block;
    k.ppl = c.ppl;
    k.gpw = c.gen_per_wave; k.ntiles = c.ntiles; k.nsys = nsys; k.ev_cap = c.ev_cap; k.req_cap = c.req_cap;
    k.park_cap = c.park_cap; k.cap = c.cap; k.slot = c.slot_doubles;
    k.n = c.n; k.npart = c.npart; k.gen = c.upload_gen;
    const void* ptrs[11] = {c.dsys, c.dtiles, c.tab, c.dev, c.dreq, c.dpark, c.x, c.tcur, c.dslow, c.dleave, c.dcnt};
    for (int q = 0; q < 11; q++) k.ptr[q] = ptrs[q];
    hipLaunchKernelGGL(set_clocks, dim3(1), dim3(256), 0, c.sm, c.tcur, nsys, time);
    HIPCHECK(hipGetLastError());
    if (!(graphs.valid && graphs.key == k)) {
        graphs.drop();
        HIPCHECK(hipStreamSynchronize(c.sm));
        HIPCHECK(hipStreamBeginCapture(c.sm, hipStreamCaptureModeThreadLocal));
        try {
            mvs_enqueue(time, h, nsteps, reset, screen, false, true);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(c.sm, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIPCHECK(hipStreamEndCapture(c.sm, &graphs.graph));
        HIPCHECK(hipGraphInstantiate(&graphs.exec, graphs.graph, nullptr, nullptr, 0));
        graphs.key = k;
        graphs.valid = true;
        graphs.captures++;
    } else {
        graphs.replays++;
    }
    HIPCHECK(hipGraphLaunch(graphs.exec, c.sm));
    HIPCHECK(hipStreamSynchronize(c.sm));
}

// ---- RA15 (algorithm 4) -----------------------------------------------------

// mdt_ra15's constants (mercury6_2.for:4162-4199): Gauss-Radau spacings h,
// the position/velocity finishing weights xc, vc, and the r, c, d tables
// built by the Fortran's recurrences in its order.
RaConst ra_constants() { return make_ra_constants(); }

// (speed of light in AU/day)^2 of mfo_grav's 1PN term: C = 299792458 m/s,
// AU = 1.495978707e11 m (mercury.inc constants, as the Fortran evaluates them)
double light_speed2() {
    const double cl = 299792458.0 * 86400.0 / (1.495978707e13 * .01);
    return cl * cl;
}

// RA15 arrays sized to the body capacity; true when they were (re)allocated,
// i.e. no RA15 history survives.
bool ensure_radau() {
    Context& c = ctx;
    if (c.ra_cap == c.cap && c.rb) return false;
    dalloc(c.rb, (size_t)21 * c.cap);
    dalloc(c.re, (size_t)21 * c.cap);
    dalloc(c.ra1, (size_t)3 * c.cap);
    if (!c.bmax) dalloc(c.bmax, 1);
    c.ra_cap = c.cap;
    return true;
}

// One force evaluation (niter = 0: a1 only) or one RA15 iteration sweep over
// every body; returns max |b7| (0 for niter = 0).
double radau_pass(double t, int niter, bool reset) {
    Context& c = ctx;
    State st = device_state();
    static const RaConst K = ra_constants();
    RaArgs args{c.dsys, st, RaState{c.rb, c.re, c.ra1}, c.tab, c.bmax, K, t, niter, reset ? 1 : 0, c.pn,
                light_speed2()};
    if (c.ra_pend) {
        // the previous sequence's finish rides on this pass; a redo of this
        // sequence then starts from the state it wrote back
        if (niter == 0) throw std::runtime_error("internal: RA15 finish pending at a force-only pass");
        args.pend = 1;
        args.ptdid = c.ra_ptdid;
        args.pq = c.ra_pq;
        args.preset = c.ra_preset ? 1 : 0;
        c.ra_pend = false;
    }
    HIPCHECK(hipMemsetAsync(c.bmax, 0, sizeof(unsigned long long), c.stream));
    const bool ng = c.ngflag != 0;
    if (c.pn && ng) hipLaunchKernelGGL((ra_planets<true, true>), dim3(1), dim3(64), 0, c.stream, args);
    else if (c.pn) hipLaunchKernelGGL((ra_planets<true, false>), dim3(1), dim3(64), 0, c.stream, args);
    else if (ng) hipLaunchKernelGGL((ra_planets<false, true>), dim3(1), dim3(64), 0, c.stream, args);
    else hipLaunchKernelGGL((ra_planets<false, false>), dim3(1), dim3(64), 0, c.stream, args);
    const int npart = c.n - c.nbig;
    if (npart > 0) {
        dim3 grid((npart + 127) / 128);
        hipEvent_t t0 = nullptr;      // kernel timing: the particle kernel counts as "particle"
        timed_launch_begin(c.stream, t0);
        if (c.pn && ng) hipLaunchKernelGGL((ra_particles<true, true>), grid, dim3(128), 0, c.stream, args);
        else if (c.pn) hipLaunchKernelGGL((ra_particles<true, false>), grid, dim3(128), 0, c.stream, args);
        else if (ng) hipLaunchKernelGGL((ra_particles<false, true>), grid, dim3(128), 0, c.stream, args);
        else hipLaunchKernelGGL((ra_particles<false, false>), grid, dim3(128), 0, c.stream, args);
        timed_launch_end(K_PARTICLE, c.stream, t0);
    }
    HIPCHECK(hipGetLastError());
    unsigned long long bits = 0;
    HIPCHECK(hipMemcpyAsync(&bits, c.bmax, sizeof(bits), hipMemcpyDeviceToHost, c.stream));
    HIPCHECK(hipStreamSynchronize(c.stream));
    double v;
    std::memcpy(&v, &bits, sizeof(v));
    return v;
}

void check_fault(const char* what);

// mdt_ra15's sequence with its step-size law on the host (libm pow, as the
// CPU path): shrink-and-redo while a reset sequence would shrink, cap 1.4x.
// The iteration runs on the device; only max |b7| comes back per pass.
void radau_step(double time, double* h, double* hdid, double tol, int64_t* forces, int64_t* rejected) {
    Context& c = ctx;
    if (c.systems.size() != 1) throw std::runtime_error("RADAU needs one system");
    if (ensure_radau()) c.history = false;
    build_systems(kRaEval, true, *h, 0.0);
    const bool reset = !c.history;
    const int niter = reset ? 6 : 2;
    double t = *h, tdid = t;
    for (;;) {
        if (time + t == time) throw std::runtime_error("RADAU step cannot advance time");
        double temp = radau_pass(t, niter, reset);
        *forces += 1 + 7 * niter;
        check_fault("Nonfinite RADAU force or predictor");
        temp = temp / (72.0 * std::pow(std::fabs(t), 7.0));
        tdid = t;
        if (temp == 0.0) t = tdid * 1.4;
        else t = std::copysign(std::pow(tol / temp, 1.0 / 9.0), tdid);
        if (reset && std::fabs(t / tdid) < 1.0) {
            t = t * .8;
            ++*rejected;
            continue;
        }
        if (std::fabs(t / tdid) > 1.4) t = tdid * 1.4;
        break;
    }
    // the finish (x, v at the end; e, b predicted for the next sequence) waits
    // for the next sequence's pass, or for radau_flush if the state is read first
    c.ra_pend = true;
    c.ra_ptdid = tdid;
    c.ra_pq = t / tdid;
    c.ra_preset = reset;
    *hdid = tdid;
    *h = t;
    c.history = true;
}

// Applies a pending RA15 finish (ra_finish over every body j >= 1: the
// step-start state kept in ox, ov; x, v at the end; e, b predicted).
void radau_flush() {
    Context& c = ctx;
    if (!c.ra_pend) return;
    c.ra_pend = false;
    static const RaConst K = ra_constants();
    RaFinishArgs fa{c.dsys, device_state(), RaState{c.rb, c.re, c.ra1}, K, c.ra_ptdid, c.ra_pq, c.ra_preset ? 1 : 0};
    if (c.n > 1) hipLaunchKernelGGL(ra_finish, dim3((c.n - 1 + 127) / 128), dim3(128), 0, c.stream, fa);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(c.stream));
}

// ---- Bulirsch-Stoer over the whole system (algorithms 2 and 3) ------------

int read_fault();
void clear_fault();
void check_ngf_fault(int f);

// The kernels of one pass over the levels LO..HI.
template <int ALG, int LO, int HI>
void bs_pass(const BswArgs& A) {
    Context& c = ctx;
    const bool ng = c.ngflag != 0;
    // the planets: kBsLanes lanes per level for up to 16 planets, kBsWide beyond
    if (c.nbig - 1 <= kBsLanes) {
        const dim3 bb((HI - LO + 1) * kBsLanes);
        if (c.pn && ng) hipLaunchKernelGGL((bsw_big<ALG, true, true, LO, HI, kBsLanes>), dim3(1), bb, 0, c.stream, A);
        else if (c.pn) hipLaunchKernelGGL((bsw_big<ALG, true, false, LO, HI, kBsLanes>), dim3(1), bb, 0, c.stream, A);
        else if (ng) hipLaunchKernelGGL((bsw_big<ALG, false, true, LO, HI, kBsLanes>), dim3(1), bb, 0, c.stream, A);
        else hipLaunchKernelGGL((bsw_big<ALG, false, false, LO, HI, kBsLanes>), dim3(1), bb, 0, c.stream, A);
    } else {
        const dim3 bb((HI - LO + 1) * kBsWide);
        if (c.pn && ng) hipLaunchKernelGGL((bsw_big<ALG, true, true, LO, HI, kBsWide>), dim3(1), bb, 0, c.stream, A);
        else if (c.pn) hipLaunchKernelGGL((bsw_big<ALG, true, false, LO, HI, kBsWide>), dim3(1), bb, 0, c.stream, A);
        else if (ng) hipLaunchKernelGGL((bsw_big<ALG, false, true, LO, HI, kBsWide>), dim3(1), bb, 0, c.stream, A);
        else hipLaunchKernelGGL((bsw_big<ALG, false, false, LO, HI, kBsWide>), dim3(1), bb, 0, c.stream, A);
    }
    const int npart = c.n - c.nbig;
    if (npart > 0) {
        dim3 grid((npart + kBswSlots - 1) / kBswSlots);
        const dim3 pb(bsw_threads<LO, HI>());
        if (c.pn && ng) hipLaunchKernelGGL((bsw_particles<ALG, true, true, LO, HI>), grid, pb, 0, c.stream, A);
        else if (c.pn) hipLaunchKernelGGL((bsw_particles<ALG, true, false, LO, HI>), grid, pb, 0, c.stream, A);
        else if (ng) hipLaunchKernelGGL((bsw_particles<ALG, false, true, LO, HI>), grid, pb, 0, c.stream, A);
        else hipLaunchKernelGGL((bsw_particles<ALG, false, false, LO, HI>), grid, pb, 0, c.stream, A);
    }
    HIPCHECK(hipGetLastError());
}

// The first level n > 3 of LO..HI whose largest error is within tol^2, 0 if none.
int bs_first_level(int lo, int hi, double tol2) {
    Context& c = ctx;
    unsigned long long bits[12];
    HIPCHECK(hipMemcpyAsync(bits, c.bserr, sizeof(bits), hipMemcpyDeviceToHost, c.stream));
    HIPCHECK(hipStreamSynchronize(c.stream));
    const int f = read_fault();
    check_ngf_fault(f);
    if (f & FAULT_FORCE) throw std::runtime_error("Nonfinite BS extrapolation");
    for (int n = std::max(lo, 4); n <= hi; n++) {
        double err;
        std::memcpy(&err, &bits[n - 1], sizeof(err));
        if (err <= tol2) return n;
    }
    return 0;
}

// One attempt of mdt_bs1 (ALG 2) or mdt_bs2 (ALG 3) at step h0: the first
// level n > 3 within tol^2, 0 if none.  The levels are independent up to the
// extrapolation, whose value at level n needs levels 1..n only, so the device
// runs levels 1..6 (BS) or 1..8 (BS2) first and the remaining ones only when
// none of those converged.  The step control (shrink after level 8, grow
// before level 7) keeps BS2's accepted level mostly within its first pass;
// BS's steady state accepts level 7, so most BS attempts run both passes --
// still faster than one pass of all 8 levels on 1M particles (6.5e8 against
// 6.2e8 particle-steps/s), since each pass pairs its levels per thread.  The
// sequential reference stops at the same first level.
template <int ALG>
int bs_attempt(double h0, double tol2) {
    Context& c = ctx;
    constexpr int L = BsShape<ALG>::L;
    constexpr int LA = ALG == 2 ? 6 : 8;
    BswArgs A{c.dsys, device_state(), c.tab, c.bscand, c.bsrsave, c.bserr, h0, light_speed2()};
    hipLaunchKernelGGL(zero_bytes, dim3(1), dim3(128), 0, c.stream, reinterpret_cast<unsigned char*>(c.bserr),
                       (long long)(12 * sizeof(unsigned long long)));
    bs_pass<ALG, 1, LA>(A);
    if (const int n = bs_first_level(1, LA, tol2)) return n;
    bs_pass<ALG, LA + 1, L>(A);
    return bs_first_level(LA + 1, L, tol2);
}

// mdt_bs1 / mdt_bs2 with mfo_all: attempts at h, halving until a level
// converges; then the Fortran's step recommendation (x .55 after the last
// level, x 1.3 before level 7).  The levels run on the device; the host reads
// the level errors after each pass of an attempt.
void bs_step(int alg, double time, double* h, double* hdid, double tol, int64_t* forces, int64_t* rejected) {
    Context& c = ctx;
    const double SHRINK = .55, GROW = 1.3;
    if (c.systems.size() != 1) throw std::runtime_error("BS needs one system");
    if (c.nbig - 1 > kBsWide) throw std::runtime_error("more than 64 Big bodies besides the central body");
    const int P = alg == 2 ? BsShape<2>::P : BsShape<3>::P;
    build_systems(1 + P, true, *h, 0.0);
    if (c.bs_cap != c.cap) {
        dalloc(c.bscand, (size_t)9 * 6 * c.cap);       // levels 4..12
        dalloc(c.bsrsave, (size_t)8 * 6 * c.cap);      // the first pass's level results
        if (!c.bserr) dalloc(c.bserr, 12);
        c.bs_cap = c.cap;
    }
    const double tol2 = tol * tol;
    *forces = 1;
    for (;;) {
        if (time + *h == time) throw std::runtime_error("BS cannot advance time at the requested tolerance");
        clear_fault();
        const int level = alg == 2 ? bs_attempt<2>(*h, tol2) : bs_attempt<3>(*h, tol2);
        if (level) {
            for (int n = 1; n <= level; n++) *forces += alg == 2 ? 2 * n : n;
            BswArgs A{c.dsys, device_state(), c.tab, c.bscand, c.bsrsave, c.bserr, *h, light_speed2()};
            if (c.n > 1) hipLaunchKernelGGL(bsw_accept, dim3((c.n - 1 + 127) / 128), dim3(128), 0, c.stream, A, level);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipStreamSynchronize(c.stream));
            *hdid = *h;
            if (level >= 8) *h = *h * SHRINK;
            if (level < 7) *h = *h * GROW;
            return;
        }
        *forces += P;
        ++*rejected;
        *h = *h * .5;
    }
}

// ---- differentiation tape -------------------------------------------------

// Planet rows (Big bodies 1..nbig-1) of x, v, a <-> psave [9][np].
void planet_rows_copy(bool save) {
    Context& c = ctx;
    const int np = c.nbig - 1;
    if (np < 1) return;
    double* arr[3] = {c.x, c.v, c.a};
    for (int a = 0; a < 3; a++)
        for (int k = 0; k < 3; k++) {
            double* dev = arr[a] + (long long)k * c.cap + 1;
            double* sv = c.tape.psave + (long long)(3 * a + k) * np;
            HIPCHECK(hipMemcpyAsync(save ? sv : dev, save ? dev : sv, sizeof(double) * np,
                                    hipMemcpyDeviceToDevice, c.stream));
        }
}

AdjArgs adj_args(int k0, int k1) {
    Context::Tape& t = ctx.tape;
    AdjArgs a{};
    a.sys = t.dsys;
    a.st = device_state();
    a.tab = t.tab;
    a.steps = t.steps;
    a.k0 = k0;
    a.k1 = k1;
    a.npart = t.npart;
    return a;
}

double* ck_at(int j) { return ctx.tape.ck + (long long)j * 6 * ctx.tape.npart; }

void adjoint_forward(int nsteps, double h, int seg, double* xo, double* vo) {
    Context& c = ctx;
    Context::Tape& t = c.tape;
    if (c.algorithm != 1 && c.algorithm != 9) throw std::runtime_error("differentiation needs the MVS algorithm");
    if (c.systems.size() != 1) throw std::runtime_error("differentiation needs one system");
    if (c.jcen[0] != 0 || c.jcen[1] != 0 || c.jcen[2] != 0)
        throw std::runtime_error("differentiation with oblateness is not provided");
    if (c.nbig - 1 > 12) throw std::runtime_error("differentiation supports at most 12 Big bodies");
    if (nsteps < 1 || !(h != 0) || !std::isfinite(h)) throw std::runtime_error("invalid step count or size");
    t.valid = false;
    const int np = c.nbig - 1;
    const int npart = c.n - c.nbig;
    const int K = seg > 0 ? std::min(seg, nsteps) : (int)std::ceil(std::sqrt((double)nsteps));
    const int nseg = (nsteps + K - 1) / K;
    const int R = row_len(np);
    Sys d{};
    d.base = 0; d.n = c.n; d.nbig = c.nbig; d.ts = table_stride(np);
    d.tab = 0; d.start = (long long)nsteps * d.ts;
    d.mu = c.systems[0].mu; d.h = h; d.hby2 = .5 * h;
    ensure_stream();
    ensure_systems(1);
    // wait for anything still running from an earlier tape
    HIPCHECK(hipStreamSynchronize(c.sg));
    dfree(t.tab); dfree(t.ck); dfree(t.rec); dfree(t.lam); dfree(t.part); dfree(t.src);
    dfree(t.grad); dfree(t.rows); dfree(t.tan); dfree(t.fin); dfree(t.seed);
    if (!t.dsys) dalloc(t.dsys, 1);
    if (!t.psave) dalloc(t.psave, 9 * kMaxBig);
    dalloc(t.tab, (size_t)(nsteps + 1) * d.ts);
    dalloc(t.ck, (size_t)(nseg + 1) * 6 * std::max(npart, 1));
    if (np > 0) dalloc(t.rows, (size_t)7 * np * (nsteps + 1) * R);
    HIPCHECK(hipMemcpy(t.dsys, &d, sizeof(Sys), hipMemcpyHostToDevice));
    t.hs = d; t.steps = nsteps; t.seg = K; t.nseg = nseg; t.npart = npart; t.np = np;
    t.nblk = (npart + 127) / 128;
    planet_rows_copy(true);
    HIPCHECK(hipEventRecord(c.ev_plan[0], c.stream));
    // the planets' tangent rows of every unit direction, on their own stream
    if (np > 0) {
        HIPCHECK(hipStreamWaitEvent(c.sg, c.ev_plan[0], 0));
        dfree(t.pfin);
        dalloc(t.pfin, (size_t)42 * np * np);
        TlmArgs ta{t.dsys, c.m, t.psave, t.psave + 3 * np, nsteps, nullptr, t.rows, t.pfin};
        hipLaunchKernelGGL(tlm_planets, dim3(7 * np), dim3(64), 0, c.sg, ta);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipEventRecord(c.ev_done[0], c.sg));
    }
    // planets: the tables, one launch per checkpoint segment, on the planet
    // stream; the particle segment j runs once tables up to its end exist
    HIPCHECK(hipStreamWaitEvent(c.sp, c.ev_plan[0], 0));
    State st = device_state();
    EventSink es{c.dev, c.devcount, c.ev_cap, c.fault};
    for (int j = 0; j < nseg; j++) {
        const int k0 = j * K, steps = std::min(nsteps, (j + 1) * K) - k0;
        PlanetArgs pa{t.dsys, st, t.tab + (long long)k0 * d.ts, c.tcur, es, 0.0, 1, steps, k0 == 0 ? 1 : 0,
                      0, 0, nullptr, 0, k0};
        hipLaunchKernelGGL(mvs_planets, dim3(1), dim3(64), 0, c.sp, pa);
        HIPCHECK(hipGetLastError());
        hipEvent_t ev = c.ev_fast[j % kRing];
        HIPCHECK(hipEventRecord(ev, c.sp));
        if (npart > 0) {
            HIPCHECK(hipStreamWaitEvent(c.stream, ev, 0));
            AdjArgs a = adj_args(k0, k0 + steps);
            a.ck_in = j == 0 ? nullptr : ck_at(j);
            a.ck_out = ck_at(j + 1);
            hipLaunchKernelGGL(adj_forward, dim3(t.nblk), dim3(128), 0, c.stream, a);
            HIPCHECK(hipGetLastError());
        }
    }
    HIPCHECK(hipStreamSynchronize(c.sp));
    // final planet states before the restore
    t.xfin.assign((size_t)3 * c.n, 0.0);
    t.vfin.assign((size_t)3 * c.n, 0.0);
    std::vector<double> row((size_t)std::max(np, 1));
    for (int k = 0; k < 3 && np > 0; k++) {
        HIPCHECK(hipMemcpy(row.data(), c.x + (long long)k * c.cap + 1, sizeof(double) * np, hipMemcpyDeviceToHost));
        for (int p = 0; p < np; p++) t.xfin[3 * (p + 1) + k] = row[p];
        HIPCHECK(hipMemcpy(row.data(), c.v + (long long)k * c.cap + 1, sizeof(double) * np, hipMemcpyDeviceToHost));
        for (int p = 0; p < np; p++) t.vfin[3 * (p + 1) + k] = row[p];
    }
    planet_rows_copy(false);
    if (npart > 0) {
        std::vector<double> fin((size_t)6 * npart);
        HIPCHECK(hipMemcpyAsync(fin.data(), ck_at(nseg), sizeof(double) * 6 * npart, hipMemcpyDeviceToHost,
                                c.stream));
        HIPCHECK(hipStreamSynchronize(c.stream));
        for (int i = 0; i < npart; i++)
            for (int k = 0; k < 3; k++) {
                t.xfin[3 * (c.nbig + i) + k] = fin[(size_t)k * npart + i];
                t.vfin[3 * (c.nbig + i) + k] = fin[(size_t)(3 + k) * npart + i];
            }
    }
    HIPCHECK(hipStreamSynchronize(c.stream));
    std::memcpy(xo, t.xfin.data(), sizeof(double) * 3 * c.n);
    std::memcpy(vo, t.vfin.data(), sizeof(double) * 3 * c.n);
    t.valid = true;
}

void require_tape() {
    if (!ctx.tape.valid) throw std::runtime_error("no tape: call mercury_hip_adjoint_forward first");
}

void adjoint_backward(const double* gx, const double* gv, double* gx0, double* gv0, double* gm) {
    Context& c = ctx;
    Context::Tape& t = c.tape;
    require_tape();
    const int np = t.np, npart = t.npart, R = row_len(np), T = t.steps;
    dfree(t.src); dfree(t.grad);
    dalloc(t.src, (size_t)(T + 1) * R);
    dalloc(t.grad, (size_t)std::max(7 * np, 1));
    std::vector<double> lam0((size_t)6 * std::max(npart, 1));
    if (npart > 0) {
        if (!t.lam) dalloc(t.lam, (size_t)6 * npart);
        if (!t.rec) dalloc(t.rec, (size_t)t.seg * 7 * npart);
        if (!t.part) dalloc(t.part, (size_t)(T + 1) * t.nblk * R);
        if (!t.fin) dalloc(t.fin, (size_t)6 * npart);
        for (int i = 0; i < npart; i++)
            for (int k = 0; k < 3; k++) {
                lam0[(size_t)k * npart + i] = gx[3 * (c.nbig + i) + k];
                lam0[(size_t)(3 + k) * npart + i] = gv[3 * (c.nbig + i) + k];
            }
        HIPCHECK(hipMemcpyAsync(t.lam, lam0.data(), sizeof(double) * 6 * npart, hipMemcpyHostToDevice, c.stream));
        for (int j = t.nseg - 1; j >= 0; j--) {
            const int k0 = j * t.seg, k1 = std::min(T, (j + 1) * t.seg);
            AdjArgs a = adj_args(k0, k1);
            a.ck_in = j == 0 ? nullptr : ck_at(j);
            a.rec = t.rec;
            hipLaunchKernelGGL(adj_forward, dim3(t.nblk), dim3(128), 0, c.stream, a);
            a.lam = t.lam;
            a.xfin = ck_at(t.nseg);
            a.fin_ell = t.fin;
            hipLaunchKernelGGL(adj_reverse, dim3(t.nblk), dim3(128), 0, c.stream, a);
            a.part = t.part;
            hipLaunchKernelGGL(adj_rows, dim3(k1 - k0 + (k1 == T ? 1 : 0), t.nblk), dim3(128), 0, c.stream, a);
            HIPCHECK(hipGetLastError());
        }
        long long tot = (long long)(T + 1) * R;
        hipLaunchKernelGGL(adj_sum_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c.stream, t.part, t.src,
                           T + 1, t.nblk, R);
        HIPCHECK(hipGetLastError());
    } else {
        HIPCHECK(hipMemsetAsync(t.src, 0, sizeof(double) * (T + 1) * R, c.stream));
    }
    std::vector<double> grad((size_t)std::max(7 * np, 1), 0.0);
    std::vector<double> pfin((size_t)42 * np * np);
    if (np > 0) {
        HIPCHECK(hipStreamWaitEvent(c.stream, c.ev_done[0], 0));
        hipLaunchKernelGGL(adj_contract, dim3(7 * np), dim3(256), 0, c.stream, t.src, t.rows, t.grad, T + 1, R);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(grad.data(), t.grad, sizeof(double) * 7 * np, hipMemcpyDeviceToHost, c.stream));
        HIPCHECK(hipMemcpyAsync(pfin.data(), t.pfin, sizeof(double) * 42 * np * np, hipMemcpyDeviceToHost,
                                c.stream));
    }
    if (npart > 0)
        HIPCHECK(hipMemcpyAsync(lam0.data(), t.lam, sizeof(double) * 6 * npart, hipMemcpyDeviceToHost, c.stream));
    HIPCHECK(hipStreamSynchronize(c.stream));
    // the planets' own final states: each direction's final tangent against
    // their cotangents (the central body's heliocentric state is identically 0)
    for (int dir = 0; dir < 7 * np; dir++) {
        const double* f = pfin.data() + (size_t)dir * 6 * np;
        double acc = 0.0;
        for (int p = 0; p < np; p++)
            for (int k = 0; k < 3; k++)
                acc += gx[3 * (p + 1) + k] * f[3 * p + k] + gv[3 * (p + 1) + k] * f[3 * np + 3 * p + k];
        grad[dir] += acc;
    }
    std::fill(gx0, gx0 + 3 * c.n, 0.0);
    std::fill(gv0, gv0 + 3 * c.n, 0.0);
    std::fill(gm, gm + c.n, 0.0);
    for (int p = 0; p < np; p++) {
        for (int k = 0; k < 3; k++) {
            gx0[3 * (p + 1) + k] = grad[3 * p + k];
            gv0[3 * (p + 1) + k] = grad[3 * np + 3 * p + k];
        }
        gm[p + 1] = grad[6 * np + p];
    }
    for (int i = 0; i < npart; i++)
        for (int k = 0; k < 3; k++) {
            gx0[3 * (c.nbig + i) + k] = lam0[(size_t)k * npart + i];
            gv0[3 * (c.nbig + i) + k] = lam0[(size_t)(3 + k) * npart + i];
        }
}

void tangent_pass(const double* dx0, const double* dv0, const double* dm, double* dx, double* dv) {
    Context& c = ctx;
    Context::Tape& t = c.tape;
    require_tape();
    const int np = t.np, npart = t.npart, R = row_len(np), T = t.steps;
    std::fill(dx, dx + 3 * c.n, 0.0);
    std::fill(dv, dv + 3 * c.n, 0.0);
    double* rows1 = nullptr;
    dalloc(rows1, (size_t)(T + 1) * R);
    try {
        if (np > 0) {
            std::vector<double> seed((size_t)7 * np);
            for (int p = 0; p < np; p++) {
                for (int k = 0; k < 3; k++) {
                    seed[3 * p + k] = dx0[3 * (p + 1) + k];
                    seed[3 * np + 3 * p + k] = dv0[3 * (p + 1) + k];
                }
                seed[6 * np + p] = dm[p + 1];
            }
            if (!t.seed) dalloc(t.seed, 7 * kMaxBig);
            double* fin = nullptr;
            dalloc(fin, 6 * kMaxBig);
            HIPCHECK(hipMemcpyAsync(t.seed, seed.data(), sizeof(double) * 7 * np, hipMemcpyHostToDevice, c.stream));
            TlmArgs ta{t.dsys, c.m, t.psave, t.psave + 3 * np, T, t.seed, rows1, fin};
            hipLaunchKernelGGL(tlm_planets, dim3(1), dim3(64), 0, c.stream, ta);
            HIPCHECK(hipGetLastError());
            std::vector<double> fh((size_t)6 * np);
            HIPCHECK(hipMemcpyAsync(fh.data(), fin, sizeof(double) * 6 * np, hipMemcpyDeviceToHost, c.stream));
            HIPCHECK(hipStreamSynchronize(c.stream));
            dfree(fin);
            for (int p = 0; p < np; p++)
                for (int k = 0; k < 3; k++) {
                    dx[3 * (p + 1) + k] = fh[3 * p + k];
                    dv[3 * (p + 1) + k] = fh[3 * np + 3 * p + k];
                }
        } else {
            HIPCHECK(hipMemsetAsync(rows1, 0, sizeof(double) * (T + 1) * R, c.stream));
        }
        if (npart > 0) {
            std::vector<double> tv((size_t)6 * npart);
            for (int i = 0; i < npart; i++)
                for (int k = 0; k < 3; k++) {
                    tv[(size_t)k * npart + i] = dx0[3 * (c.nbig + i) + k];
                    tv[(size_t)(3 + k) * npart + i] = dv0[3 * (c.nbig + i) + k];
                }
            if (!t.tan) dalloc(t.tan, (size_t)6 * npart);
            HIPCHECK(hipMemcpyAsync(t.tan, tv.data(), sizeof(double) * 6 * npart, hipMemcpyHostToDevice, c.stream));
            AdjArgs a = adj_args(0, T);
            a.rows = rows1;
            a.tan = t.tan;
            hipLaunchKernelGGL(tlm_particles, dim3(t.nblk), dim3(128), 0, c.stream, a);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipMemcpyAsync(tv.data(), t.tan, sizeof(double) * 6 * npart, hipMemcpyDeviceToHost, c.stream));
            HIPCHECK(hipStreamSynchronize(c.stream));
            for (int i = 0; i < npart; i++)
                for (int k = 0; k < 3; k++) {
                    dx[3 * (c.nbig + i) + k] = tv[(size_t)k * npart + i];
                    dv[3 * (c.nbig + i) + k] = tv[(size_t)(3 + k) * npart + i];
                }
        }
    } catch (...) {
        dfree(rows1);
        throw;
    }
    dfree(rows1);
}

// ---- HYBRID -----------------------------------------------------------------

int read_fault();
void clear_fault();

void ensure_hybrid() {
    Context& c = ctx;
    if (c.hy_cap != c.cap || !c.rcrit) {
        dalloc(c.rcrit, (size_t)c.cap);
        dalloc(c.hx0, (size_t)3 * c.cap);
        dalloc(c.hv0, (size_t)3 * c.cap);
        HIPCHECK(hipMemset(c.rcrit, 0, sizeof(double) * c.cap));
        c.hy_cap = c.cap;
    }
    if (!c.hytab) dalloc(c.hytab, (size_t)(kHyHdr + HB_N * kMaxBig));
    if (!c.hnpairs) dalloc(c.hnpairs, 1);
    // room for every possible pair (each particle with every Big body), so
    // the scan can never overflow after the state has moved
    const long long np = std::max(c.nbig - 1, 1);
    const long long need = std::max<long long>(1 << 16, (long long)(c.n - c.nbig) * np + np * np);
    if (need > (1LL << 31) - 1) throw std::runtime_error("too many possible HYBRID pairs");
    if (c.hpair_cap < need) {
        c.hpair_cap = (int)need;
        dalloc(c.hpairs, (size_t)2 * c.hpair_cap);
    }
}

HyArgs hy_args(bool reset) {
    Context& c = ctx;
    HyArgs a{};
    a.sys = c.dsys;
    a.st = device_state();
    a.rcrit = c.rcrit;
    a.x0 = c.hx0;
    a.v0 = c.hv0;
    a.tab = c.hytab;
    a.pairs = c.hpairs;
    a.npairs = c.hnpairs;
    a.cap = c.hpair_cap;
    a.reset = reset ? 1 : 0;
    a.fault = c.fault;
    return a;
}

// mdt_hy's first half and the changeover scan (mercury6_2.for:3655-3727):
// kick, Sun-momentum drift, Kepler drift of every body, mce_snif.  Encounter
// members go back to their pre-drift states, which the host receives.
void hybrid_begin(double h, const double* crit, int flag, int cap, int* ce, int* nce, int* pi, int* pj, double* xx,
                  double* vv, int64_t* forces) {
    Context& c = ctx;
    if (c.algorithm != 10) throw std::runtime_error("HYBRID calls need algorithm 10");
    if (c.enc.active) throw std::runtime_error("HYBRID step inside an encounter context");
    if (c.systems.size() != 1) throw std::runtime_error("HYBRID needs one system");
    if (!(h != 0) || !std::isfinite(h)) throw std::runtime_error("invalid step");
    ensure_stream();
    ensure_hybrid();
    build_systems(1, true, h, 0.0);
    const bool reset = flag != 2 || !c.history;
    if (reset) HIPCHECK(hipMemcpy(c.rcrit, crit, sizeof(double) * c.n, hipMemcpyHostToDevice));
    *forces = reset ? 1 : 0;
    c.tape.valid = false;
    for (;;) {
        clear_fault();
        HIPCHECK(hipMemsetAsync(c.hnpairs, 0, sizeof(int), c.stream));
        HyArgs a = hy_args(reset);
        // a reset recomputes accelerations from the start state: run on a copy
        // of x, v when the pair buffer has to grow and the step is retried
        hipLaunchKernelGGL(hy_big_begin, dim3(1), dim3(64), 0, c.stream, a);
        const int npart = c.n - c.nbig;
        const bool ng = c.ngflag == 1 || c.ngflag == 3;
        if (npart > 0 && ng)
            hipLaunchKernelGGL(hy_particles_begin<true>, dim3((npart + 127) / 128), dim3(128), 0, c.stream, a);
        else if (npart > 0)
            hipLaunchKernelGGL(hy_particles_begin<false>, dim3((npart + 127) / 128), dim3(128), 0, c.stream, a);
        HIPCHECK(hipGetLastError());
        int cnt = 0;
        HIPCHECK(hipMemcpyAsync(&cnt, c.hnpairs, sizeof(int), hipMemcpyDeviceToHost, c.stream));
        HIPCHECK(hipStreamSynchronize(c.stream));
        const int f = read_fault();
        check_ngf_fault(f);
        if (f & FAULT_FORCE) throw std::runtime_error("HYBRID force failed");
        if (cnt > c.hpair_cap) {
            // the scan overflowed after the state moved: the encounter list is
            // incomplete, and a retry would double-step; fail loudly
            throw std::runtime_error("HYBRID encounter pair buffer overflow");
        }
        std::vector<int> pr((size_t)2 * cnt);
        if (cnt) HIPCHECK(hipMemcpy(pr.data(), c.hpairs, sizeof(int) * 2 * cnt, hipMemcpyDeviceToHost));
        std::vector<std::pair<int, int>> pairs((size_t)cnt);
        for (int k = 0; k < cnt; k++) pairs[k] = {pr[2 * k], pr[2 * k + 1]};
        std::sort(pairs.begin(), pairs.end());
        if (cnt > cap) throw std::runtime_error("Hybrid encounter capacity exceeded");
        *nce = cnt;
        std::fill(ce, ce + c.n, 0);
        for (int k = 0; k < cnt; k++) {
            pi[k] = pairs[k].first + 1;
            pj[k] = pairs[k].second + 1;
            ce[pairs[k].first] = 2;
            ce[pairs[k].second] = 2;
        }
        c.members.clear();
        for (int j = 1; j < c.n; j++)
            if (ce[j]) c.members.push_back(j);
        const int ns = (int)c.members.size();
        if (ns) {
            if (ns > c.hpack_cap) {
                c.hpack_cap = std::max(ns, 2 * c.hpack_cap);
                dalloc(c.hpack, (size_t)6 * c.hpack_cap);
                dalloc(c.hids, (size_t)c.hpack_cap);
            }
            HIPCHECK(hipMemcpy(c.hids, c.members.data(), sizeof(int) * ns, hipMemcpyHostToDevice));
            hipLaunchKernelGGL(hy_revert_gather, dim3((ns + 127) / 128), dim3(128), 0, c.stream, device_state(),
                               c.hx0, c.hv0, c.hids, ns, c.hpack);
            HIPCHECK(hipGetLastError());
            c.packed.resize((size_t)6 * ns);
            HIPCHECK(hipMemcpyAsync(c.packed.data(), c.hpack, sizeof(double) * 6 * ns, hipMemcpyDeviceToHost,
                                    c.stream));
            HIPCHECK(hipStreamSynchronize(c.stream));
            for (int a2 = 0; a2 < ns; a2++)
                for (int k = 0; k < 3; k++) {
                    xx[3 * c.members[a2] + k] = c.packed[6 * a2 + k];
                    vv[3 * c.members[a2] + k] = c.packed[6 * a2 + 3 + k];
                }
        }
        break;
    }
    c.history = true;
    c.saved_valid = true;
}

// mdt_hy's second half (mercury6_2.for:3758-3796): the members' states from
// the host, Sun-momentum drift, mfo_hy, kick.
void hybrid_finish(double h, const double* m, const double* xx, const double* vv, int64_t* forces) {
    Context& c = ctx;
    if (c.algorithm != 10) throw std::runtime_error("HYBRID calls need algorithm 10");
    if (c.enc.active) throw std::runtime_error("HYBRID step inside an encounter context");
    build_systems(1, true, h, 0.0);
    const int ns = (int)c.members.size();
    if (ns) {
        for (int a2 = 0; a2 < ns; a2++)
            for (int k = 0; k < 3; k++) {
                c.packed[6 * a2 + k] = xx[3 * c.members[a2] + k];
                c.packed[6 * a2 + 3 + k] = vv[3 * c.members[a2] + k];
            }
        HIPCHECK(hipMemcpyAsync(c.hpack, c.packed.data(), sizeof(double) * 6 * ns, hipMemcpyHostToDevice, c.stream));
        hipLaunchKernelGGL(hy_scatter, dim3((ns + 127) / 128), dim3(128), 0, c.stream, device_state(), c.hids, ns,
                           c.hpack);
        // encounter collisions may change Big masses; Small bodies stay massless
        for (int j = c.nbig; j < c.n; j++)
            if (m[j] != 0) throw std::runtime_error("Small bodies with mass are not provided by the HIP path");
        HIPCHECK(hipMemcpyAsync(c.m, m, sizeof(double) * c.nbig, hipMemcpyHostToDevice, c.stream));
    }
    clear_fault();
    HyArgs a = hy_args(false);
    hipLaunchKernelGGL(hy_big_finish, dim3(1), dim3(64), 0, c.stream, a);
    const int npart = c.n - c.nbig;
    if (npart > 0 && (c.ngflag == 1 || c.ngflag == 3))
        hipLaunchKernelGGL(hy_particles_finish<true>, dim3((npart + 127) / 128), dim3(128), 0, c.stream, a);
    else if (npart > 0)
        hipLaunchKernelGGL(hy_particles_finish<false>, dim3((npart + 127) / 128), dim3(128), 0, c.stream, a);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(c.stream));
    check_fault("HYBRID force failed");
    c.members.clear();
    *forces = 1;
}

// ---- compact encounter context (BS2) -------------------------------------

void enc_settle();

// Each particle's partners, their max(rcrit) and masses, packed for the BS2
// kernels; again whenever the masses change.
void encounter_pack() {
    Context& c = ctx;
    Context::Enc& e = c.enc;
    if (e.n <= e.nbig) return;
    BsArgs A{};
    A.n = e.n; A.nbig = e.nbig; A.m = e.m; A.rcrit = e.rcrit; A.poff = e.poff; A.pidx = e.pidx;
    hipLaunchKernelGGL(pack_partners, dim3((e.n - e.nbig + 255) / 256), dim3(256), 0, c.stream, A, e.pinfo, e.pcon);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(c.stream));
}

void encounter_enter(int n, int nbig, const double* m, const double* x, const double* v, const double* crit,
                     int np_, const int* pi, const int* pj) {
    enc_settle();
    Context& c = ctx;
    Context::Enc& e = c.enc;
    if (c.algorithm != 10) throw std::runtime_error("encounter context needs algorithm 10");
    if (e.active) throw std::runtime_error("encounter context already active");
    if (n < 1 || nbig < 1 || nbig > n) throw std::runtime_error("invalid encounter body counts");
    if (nbig - 1 > kEncBig) throw std::runtime_error("more than 16 Big bodies in one encounter set");
    for (int j = nbig; j < n; j++)
        if (m[j] != 0) throw std::runtime_error("Small bodies with mass are not provided by the HIP path");
    ensure_stream();
    if (n > e.ncap) {
        int cap = std::max(n, 2 * e.ncap);
        for (double** p : {&e.m, &e.rcrit}) dalloc(*p, (size_t)cap);
        for (double** p : {&e.x, &e.v}) dalloc(*p, (size_t)3 * cap);
        dalloc(e.cand, (size_t)108 * cap);   // two rotating sets of 9 levels x 6 components
        dalloc(e.poff, (size_t)cap + 1);
        dalloc(e.pinfo, (size_t)cap);
        dalloc(e.pcon, (size_t)cap);
        e.ncap = cap;
    }
    if (!e.err) dalloc(e.err, 36);        // three rotating sets of 12 level errors
    if (!e.tp) dalloc(e.tp, (size_t)kBsPoints * 3 * kEncBig);
    // pair lists: Big-Big pairs in list order; each particle's Big partners
    std::vector<int> bb, off((size_t)(n - nbig) + 1, 0), idx;
    std::vector<std::vector<int>> part((size_t)std::max(n - nbig, 0));
    for (int k = 0; k < np_; k++) {
        int i = pi[k] - 1, j = pj[k] - 1;
        if (i < 1 || j < 1 || i >= n || j >= n) throw std::runtime_error("encounter pair out of range");
        if (i >= nbig) throw std::runtime_error("encounter pair without a Big body first");
        if (j < nbig) { bb.push_back(i); bb.push_back(j); }
        else part[j - nbig].push_back(i);
    }
    for (int q = 0; q < n - nbig; q++) {
        off[q + 1] = off[q] + (int)part[q].size();
        idx.insert(idx.end(), part[q].begin(), part[q].end());
    }
    if ((int)bb.size() > 2 * kBbMax) throw std::runtime_error("more than 256 Big-Big pairs in one encounter set");
    if ((int)bb.size() > e.nbbp_cap) {
        e.nbbp_cap = std::max((int)bb.size(), 64);
        dalloc(e.bbp, (size_t)e.nbbp_cap);
    }
    if ((int)idx.size() > e.npidx_cap) {
        e.npidx_cap = std::max((int)idx.size(), 64);
        dalloc(e.pidx, (size_t)e.npidx_cap);
    }
    if (!bb.empty()) HIPCHECK(hipMemcpy(e.bbp, bb.data(), sizeof(int) * bb.size(), hipMemcpyHostToDevice));
    if (!idx.empty()) HIPCHECK(hipMemcpy(e.pidx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(e.poff, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice));
    e.nbbp = (int)bb.size() / 2;
    e.n = n;
    e.nbig = nbig;
    HIPCHECK(hipMemcpy(e.m, m, sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(e.rcrit, crit, sizeof(double) * n, hipMemcpyHostToDevice));
    encounter_pack();
    e.active = true;
    // states
    e.stage.resize((size_t)3 * n);
    for (const double* src : {x, v}) {
        for (int j = 0; j < n; j++)
            for (int k = 0; k < 3; k++) e.stage[(size_t)k * n + j] = src[3 * j + k];
        double* dst = src == x ? e.x : e.v;
        for (int k = 0; k < 3; k++)
            HIPCHECK(hipMemcpy(dst + (size_t)k * e.ncap, e.stage.data() + (size_t)k * n, sizeof(double) * n,
                               hipMemcpyHostToDevice));
    }
}

void encounter_states(const double* x, const double* v) {
    enc_settle();
    Context::Enc& e = ctx.enc;
    e.stage.resize((size_t)3 * e.n);
    for (const double* src : {x, v}) {
        for (int j = 0; j < e.n; j++)
            for (int k = 0; k < 3; k++) e.stage[(size_t)k * e.n + j] = src[3 * j + k];
        double* dst = src == x ? e.x : e.v;
        for (int k = 0; k < 3; k++)
            HIPCHECK(hipMemcpy(dst + (size_t)k * e.ncap, e.stage.data() + (size_t)k * e.n, sizeof(double) * e.n,
                               hipMemcpyHostToDevice));
    }
}

// The stream's pending work on the encounter buffers (a substep's accept,
// launched without a wait) ends before the host or a null-stream copy touches
// them.
void enc_settle() {
    if (ctx.enc.pending) {
        HIPCHECK(hipStreamSynchronize(ctx.stream));
        ctx.enc.pending = false;
    }
}

// Members' states in the driver's layout: packed on the device behind the
// last substep's kernels, one copy on the stream into pinned staging, and one
// wait.
void encounter_download(double* x, double* v) {
    Context& c = ctx;
    Context::Enc& e = c.enc;
    const size_t need = (size_t)6 * e.n;
    if (e.hstage_cap < need) {
        if (e.hstage) HIPCHECK(hipHostFree(e.hstage));
        e.hstage = nullptr;
        HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&e.hstage), sizeof(double) * std::max<size_t>(need, 6),
                               hipHostMallocDefault));
        HIPCHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&e.dstage), e.hstage, 0));
        e.hstage_cap = need;
    }
    // the pack writes the driver's layout straight into the pinned staging
    hipLaunchKernelGGL(enc_pack_aos, dim3((e.n + 255) / 256), dim3(256), 0, c.stream, e.x, e.v, e.n, e.ncap,
                       e.dstage);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(c.stream));
    e.pending = false;
    std::memcpy(x, e.hstage, sizeof(double) * 3 * e.n);
    std::memcpy(v, e.hstage + (size_t)3 * e.n, sizeof(double) * 3 * e.n);
}

// mdt_bs2 (mercury6_2.for:3428-3617) over the encounter set: every attempt
// evaluates the 12 modified-midpoint levels on the device; the host takes the
// first level n > 3 whose largest error is within tol^2, or halves the step.
void bs2_step(double time, double* h, double* hdid, double tol, int64_t* forces, int64_t* rejected) {
    Context& c = ctx;
    Context::Enc& e = c.enc;
    const double SHRINK = .55, GROW = 1.3;
    const int npart = e.n - e.nbig;
    BsArgs A{};
    A.n = e.n; A.nbig = e.nbig; A.m = e.m; A.rcrit = e.rcrit; A.x = e.x; A.v = e.v; A.ncap = e.ncap;
    A.bbp = e.bbp; A.nbbp = e.nbbp; A.poff = e.poff; A.pidx = e.pidx; A.pinfo = e.pinfo; A.pcon = e.pcon;
    A.tol2 = tol * tol;
    A.tp = e.tp; A.cand = e.cand; A.err = e.err; A.fault = c.fault;
    *forces = 1;           // the start accelerations, once per call
    if (!e.hbits) {
        HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&e.hbits), 16 * sizeof(unsigned long long),
                               hipHostMallocDefault));
        HIPCHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&e.dbits), e.hbits, 0));
    }
    for (;;) {
        if (time + *h == time) throw std::runtime_error("BS cannot advance time at the requested tolerance");
        A.h0 = *h;
        // one wait per attempt: one kernel clears the fault flags and level
        // errors, one writes them into the pinned words, behind the previous
        // substep's accept on the stream
        hipLaunchKernelGGL(bs_attempt_clear, dim3(1), dim3(64), 0, c.stream, c.fault, e.err);
        hipLaunchKernelGGL(bs2_big, dim3(1), dim3(kLevels * kEncBig), 0, c.stream, A);
        if (npart > 0)
            hipLaunchKernelGGL(bs2_particles, dim3((npart + kBsPairSlots - 1) / kBsPairSlots), dim3(kBsPairThreads),
                               0, c.stream, A);
        hipLaunchKernelGGL(bs_attempt_report, dim3(1), dim3(64), 0, c.stream, e.err, c.fault, e.dbits);
        HIPCHECK(hipGetLastError());
        unsigned long long* bits = e.hbits;
        HIPCHECK(hipStreamSynchronize(c.stream));
        e.pending = false;
        int fault = 0;
        std::memcpy(&fault, bits + 12, sizeof(int));
        if (fault & FAULT_FORCE) throw std::runtime_error("Nonfinite BS extrapolation");
        int level = 0;
        for (int n = 4; n <= 12 && !level; n++) {
            double err;
            std::memcpy(&err, &bits[n - 1], sizeof(err));
            if (err <= A.tol2) level = n;
        }
        if (level) {
            for (int n = 1; n <= level; n++) *forces += n;
            if (e.n > 1) {
                hipLaunchKernelGGL(bs2_accept, dim3((e.n - 1 + 127) / 128), dim3(128), 0, c.stream, A, level);
                e.pending = true;   // settled by the next reader of the encounter buffers
            }
            HIPCHECK(hipGetLastError());
            *hdid = *h;
            if (level >= 8) *h = *h * SHRINK;
            if (level < 7) *h = *h * GROW;
            return;
        }
        *forces += 78;
        ++*rejected;
        *h = *h * .5;
    }
}

// The driver's mdt_hkce loop over one step h0 (mercury6_2.for:3883-3905) with
// the step control on the device: per attempt one bs2_loop_fused launch when
// all its blocks are resident at once, else one bs2_loop_paired launch when
// its blocks are (twice the particles per block, levels paired), else
// bs2_loop_big and bs2_loop_particles; queued in
// batches without host synchronisation (launches after the loop's end only
// carry the state forward); the host reads the state between batches.
// MERCURY_HIP_ENC_LOOP=fused, =paired or =split pins a variant (all give the same bits).
// Same sequence of steps and states as calling bs2_step per substep; the
// encounter listing the driver makes after each substep is not made.
void encounter_advance(double h0, double* hrec, double tol, int64_t* substeps, int64_t* forces,
                       int64_t* rejected) {
    enc_settle();
    Context& c = ctx;
    Context::Enc& e = c.enc;
    if (!e.active) throw std::runtime_error("no encounter context");
    if (!(h0 != 0) || !std::isfinite(h0) || !std::isfinite(*hrec) || !(tol > 0))
        throw std::runtime_error("invalid encounter step");
    if (!e.ctl) dalloc(e.ctl, 2);
    BsCtl k{};
    k.h0 = h0;
    k.tlocal = 0.0;
    k.tol2 = tol * tol;
    double hlocal = std::copysign(*hrec, h0);
    const double tmp0 = std::fabs(h0) - std::fabs(k.tlocal);
    k.hrec = hlocal;
    if (std::fabs(hlocal) > tmp0) hlocal = std::copysign(tmp0, h0);
    k.hlocal = hlocal;
    k.forces = 1;
    if (k.tlocal + k.hlocal == k.tlocal) throw std::runtime_error("BS cannot advance time at the requested tolerance");
    const int npart = e.n - e.nbig;
    BsArgs A{};
    A.n = e.n; A.nbig = e.nbig; A.m = e.m; A.rcrit = e.rcrit; A.x = e.x; A.v = e.v; A.ncap = e.ncap;
    A.bbp = e.bbp; A.nbbp = e.nbbp; A.poff = e.poff; A.pidx = e.pidx; A.pinfo = e.pinfo; A.pcon = e.pcon;
    A.tol2 = k.tol2;
    A.tp = e.tp; A.cand = e.cand; A.err = e.err; A.fault = c.fault;
    const BsLoop B{e.ctl, e.err, e.cand};
    clear_fault();
    HIPCHECK(hipMemcpyAsync(e.ctl, &k, sizeof(k), hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(zero_bytes, dim3(1), dim3(128), 0, c.stream, reinterpret_cast<unsigned char*>(e.err),
                       (long long)(36 * sizeof(unsigned long long)));
    HIPCHECK(hipGetLastError());
    const dim3 gf((unsigned)std::max((npart + kBsPart - 1) / kBsPart, 1));
    const dim3 gp((unsigned)std::max((npart + kBsPairSlots - 1) / kBsPairSlots, 1));
    int per_cu = 0, ncu = 0, dev = 0;
    HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&bs2_loop_fused),
                                                          kLevels * kBsPart, 0));
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    bool fused = (long long)gf.x <= (long long)per_cu * ncu;
    // otherwise one paired launch per attempt when its blocks are resident at once
    const size_t tb = sizeof(double) * kBsPoints * 3 * (size_t)std::max(e.nbig - 1, 0);
    int per_cu_p = 0;
    HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_p, reinterpret_cast<const void*>(&bs2_loop_paired),
                                                          kBsPairThreads, tb));
    bool paired = !fused && (long long)gp.x <= (long long)per_cu_p * ncu;
    if (const char* v = std::getenv("MERCURY_HIP_ENC_LOOP")) {
        fused = !std::strcmp(v, "fused");
        paired = !std::strcmp(v, "paired");
    }
    int q = 0;
    for (int batch = 8;; batch = std::min(2 * batch, 256)) {
        for (int j = 0; j < batch; j++, q++) {
            if (paired) {
                hipLaunchKernelGGL(bs2_loop_paired, gp, dim3(kBsPairThreads), tb, c.stream, A, B, q);
                continue;
            }
            if (fused) {
                hipLaunchKernelGGL(bs2_loop_fused, gf, dim3(kLevels * kBsPart), 0, c.stream, A, B, q);
                continue;
            }
            hipLaunchKernelGGL(bs2_loop_big, dim3(1), dim3(kLevels * kEncBig), 0, c.stream, A, B, q);
            if (npart > 0) hipLaunchKernelGGL(bs2_loop_particles, gp, dim3(kBsPairThreads), 0, c.stream, A, B, q);
        }
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(&k, e.ctl + (q & 1), sizeof(k), hipMemcpyDeviceToHost, c.stream));
        HIPCHECK(hipStreamSynchronize(c.stream));
        if (k.state == 3) throw std::runtime_error("Nonfinite BS extrapolation");
        if (k.state == 2) throw std::runtime_error("BS cannot advance time at the requested tolerance");
        if (k.state == 1) break;
    }
    *hrec = k.hrec;
    *substeps = k.substeps;
    *forces = k.forces;
    *rejected = k.rejected;
}

int read_fault() {
    int bad = 0;
    HIPCHECK(hipMemcpy(&bad, ctx.fault, sizeof(int), hipMemcpyDeviceToHost));
    return bad;
}

// mfo_ngf's fail: a comet with A2 or yar and no transverse velocity, or with
// A3 and no orbit normal
void check_ngf_fault(int f) {
    if (f & FAULT_NGF)
        throw std::runtime_error("non-gravitational force failed: undefined transverse direction or orbit normal");
}

void check_fault(const char* what) {
    int f = read_fault();
    check_ngf_fault(f);
    if (f & FAULT_FORCE) throw std::runtime_error(std::string(what) + ": non-finite acceleration");
    if (f & FAULT_REQ_OVERFLOW)
        throw std::runtime_error("screening request buffer overflow: use fewer steps per launch");
    if (f & FAULT_EVENT_OVERFLOW) throw std::runtime_error("event buffer overflow: advance fewer steps per call");
}

void clear_fault() {
    HIPCHECK(hipMemset(ctx.fault, 0, sizeof(int)));
}

void require_state() {
    if (!ctx.have_state) throw std::runtime_error("no state uploaded");
}

bool mvs_family() { return ctx.algorithm == 1 || ctx.algorithm == 9; }

MercuryEvent to_public(const DevEvent& e) {
    MercuryEvent r;
    r.i = e.i; r.j = e.j; r.kind = e.kind; r.now = e.now;
    r.distance = e.distance; r.time = e.time;
    for (int k = 0; k < 6; k++) { r.xi[k] = e.xi[k]; r.xj[k] = e.xj[k]; }
    return r;
}

bool pair_less(int ai, int aj, int bi, int bj) {
    int a0 = std::min(ai, aj), a1 = std::max(ai, aj), b0 = std::min(bi, bj), b1 = std::max(bi, bj);
    return a0 < b0 || (a0 == b0 && a1 < b1);
}

// Fetch `count` device events into hev.
void fetch_events(int count) {
    ctx.hev.resize(count);
    if (count)
        HIPCHECK(hipMemcpy(ctx.hev.data(), ctx.dev, sizeof(DevEvent) * (size_t)count, hipMemcpyDeviceToHost));
}

}  // namespace

// ===========================================================================
extern "C" {

int mercury_hip_available(void) {
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

int mercury_hip_configure(int algorithm) {
    if (algorithm < 1 || algorithm > 4) {
        if (algorithm != 9 && algorithm != 10) return 1;
    }
    if (algorithm != ctx.algorithm) release_all();
    ctx.algorithm = algorithm;
    ctx.history = false;
    return 0;
}

void mercury_hip_reset(int flag) {
    if (ctx.enc.active) return;      // BS2 keeps no history between calls
    if (flag != 2) ctx.history = false;
}

int mercury_hip_limits(const double* rce) {
    try {
        require_state();
        HIPCHECK(hipMemcpy(ctx.rce, rce, sizeof(double) * ctx.n, hipMemcpyHostToDevice));
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int mercury_hip_upload(int n, int nbig, int pn, int ngflag, const double* m, const double* x, const double* v,
                       const double* ngf, const double* jcen, const double* rce, const double* rphys) {
    try {
        if (n < 1 || nbig < 1 || nbig > n) throw std::runtime_error("invalid body counts");
        if (nbig > kMaxBig) throw std::runtime_error("more Big bodies than the HIP path supports (64)");
        if (!(m[0] > 0)) throw std::runtime_error("central mass must be positive");
        if (ngflag < 0 || ngflag > 3) throw std::runtime_error("invalid non-gravitational flag");
        if (pn != 0 && (mvs_family() || ctx.algorithm == 10))
            throw std::runtime_error("MVS and HYBRID do not take post-Newtonian forces");
        // the symplectic maps apply only mfo_ngf (ngflag 1 or 3): HYBRID on every
        // body, MVS on particles here
        const bool cometary = (ngflag == 1 || ngflag == 3);
        auto has_ngf = [&](int j) {
            return ngf[5 * j] != 0.0 || ngf[5 * j + 1] != 0.0 || ngf[5 * j + 2] != 0.0 || ngf[5 * j + 4] != 0.0;
        };
        std::vector<int> comets;
        if (cometary && mvs_family()) {
            for (int j = 1; j < n; j++) {
                if (!has_ngf(j)) continue;
                if (j < nbig)
                    throw std::runtime_error("MVS with non-gravitational forces on Big bodies is not provided by the "
                                             "HIP path");
                comets.push_back(j);
            }
        }
        const bool need_angf = !comets.empty() || (cometary && ctx.algorithm == 10);
        if (ctx.enc.active) throw std::runtime_error("upload inside an encounter context");
        for (int j = nbig; j < n; j++)
            if (m[j] != 0) throw std::runtime_error("Small bodies with mass are not provided by the HIP path");
        for (int j = 0; j < 3 * n; j++)
            if (!std::isfinite(x[j]) || !std::isfinite(v[j])) throw std::runtime_error("nonfinite state");
        ensure_stream();
        ensure_capacity(n);
        ctx.n = n; ctx.nbig = nbig; ctx.pn = pn; ctx.ngflag = ngflag;
        for (int k = 0; k < 3; k++) ctx.jcen[k] = jcen[k];
        ctx.systems.assign(1, Context::HostSys{0, n, nbig, m[0], 0.0, 0.0, jcen[0], jcen[1], jcen[2]});
        ctx.ensemble = false;
        ctx.sys_valid = false;
        ctx.upload_gen++;
        ctx.ra_pend = false;   // a new state: the last sequence's finish is moot
        if (ngflag) {
            // parameters component-major: [q][body], q = A1, A2, A3, beta, yar
            if (ctx.ngf_cap != ctx.cap) {   // the kernels index it with the body capacity
                dalloc(ctx.dngf, (size_t)5 * ctx.cap);
                ctx.ngf_cap = ctx.cap;
            }
            std::vector<double> col(n);
            for (int q = 0; q < 5; q++) {
                for (int j = 0; j < n; j++) col[j] = ngf[5 * j + q];
                HIPCHECK(hipMemcpy(ctx.dngf + (size_t)q * ctx.ngf_cap, col.data(), sizeof(double) * n,
                                   hipMemcpyHostToDevice));
            }
        }
        ctx.comets = comets;
        if (need_angf) {
            if (ctx.angf_cap != ctx.cap) {
                dalloc(ctx.dangf, (size_t)3 * ctx.cap);
                ctx.angf_cap = ctx.cap;
            }
            HIPCHECK(hipMemset(ctx.dangf, 0, sizeof(double) * 3 * ctx.angf_cap));
        }
        clear_slow_set();
        HIPCHECK(hipMemcpy(ctx.m, m, sizeof(double) * n, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(ctx.rce, rce, sizeof(double) * n, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(ctx.rphys, rphys, sizeof(double) * n, hipMemcpyHostToDevice));
        upload_vec3(x, ctx.x);
        upload_vec3(v, ctx.v);
        HIPCHECK(hipMemset(ctx.a, 0, sizeof(double) * 3 * ctx.cap));
        HIPCHECK(hipMemset(ctx.ox, 0, sizeof(double) * 3 * ctx.cap));
        HIPCHECK(hipMemset(ctx.ov, 0, sizeof(double) * 3 * ctx.cap));
        HIPCHECK(hipMemset(ctx.fault, 0, sizeof(int)));
        build_tiles();
        ctx.history = false;
        ctx.saved_valid = false;
        ctx.have_state = true;
        ctx.tape.valid = false;
        HIPCHECK(hipDeviceSynchronize());
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int mercury_hip_step(double time, double* h, double* hdid, double tol, int64_t* forces, int64_t* rejected) {
    try {
        (void)tol;
        require_state();
        *rejected = 0;
        *forces = 0;
        if (ctx.enc.active) {
            bs2_step(time, h, hdid, tol, forces, rejected);
            return 0;
        }
        if (ctx.algorithm == 10) throw std::runtime_error("HYBRID steps are mercury_hip_hybrid_begin/finish");
        if (ctx.algorithm == 4) {
            clear_fault();
            ctx.tape.valid = false;
            radau_step(time, h, hdid, tol, forces, rejected);
            ctx.saved_valid = true;
            return 0;
        }
        if (ctx.algorithm == 2 || ctx.algorithm == 3) {
            ensure_stream();
            ctx.tape.valid = false;
            bs_step(ctx.algorithm, time, h, hdid, tol, forces, rejected);
            ctx.saved_valid = true;
            return 0;
        }
        if (!mvs_family()) throw std::runtime_error("algorithm not provided by the HIP path");
        ctx.tape.valid = false;
        bool reset = !ctx.history;
        mvs_run(time, *h, 1, reset, false, true, 0.0);   // clears and reports the fault flag itself
        check_ngf_fault(ctx.hstatus[0]);
        if (ctx.hstatus[0] & FAULT_FORCE) throw std::runtime_error("MVS force failed: non-finite acceleration");
        ctx.history = true;
        ctx.saved_valid = true;
        *hdid = *h;
        *forces = reset ? 2 : 1;
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int mercury_hip_download(double* x, double* v, int previous) {
    try {
        if (ctx.enc.active) {
            if (previous) throw std::runtime_error("no previous state in an encounter context");
            encounter_download(x, v);
            return 0;
        }
        require_state();
        radau_flush();
        if (previous && !ctx.saved_valid) throw std::runtime_error("no previous state");
        download_vec3(previous ? ctx.ox : ctx.x, x);
        download_vec3(previous ? ctx.ov : ctx.v, v);
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int mercury_hip_events(double time, double h, double rcen, const MercuryEvent** out, int* count) {
    try {
        require_state();
        radau_flush();
        if (!ctx.saved_valid) throw std::runtime_error("no step to screen");
        // single-system integrators that keep the step-start states in ox, ov
        const bool ra = ctx.algorithm == 4 || ctx.algorithm == 2 || ctx.algorithm == 3;
        const bool hy = ctx.algorithm == 10;
        if (ctx.enc.active) throw std::runtime_error("event screening inside an encounter context");
        if (!mvs_family() && !ra && !hy) throw std::runtime_error("algorithm not provided by the HIP path");
        if (hy) build_systems(1, true, h, rcen);
        int n = ctx.n;
        if (ra && ctx.nbig > 1) {
            // planet boxes and end positions from the step-start and end states
            hipLaunchKernelGGL(planet_screen_block, dim3(1), dim3(64), 0, ctx.stream, ctx.dsys, device_state(),
                               ctx.tab + ctx.hsys[0].tab, h);
            HIPCHECK(hipGetLastError());
        }
        for (;;) {
            HIPCHECK(hipMemsetAsync(ctx.devcount, 0, sizeof(int), ctx.stream));
            SavedScreenArgs sa{ctx.dsys, device_state(), ctx.tab, EventSink{ctx.dev, ctx.devcount, ctx.ev_cap, ctx.fault},
                               time, h, rcen, hy ? 0 : 1, ra ? ctx.ov : ctx.v0k, ra ? ctx.v : ctx.v1,
                               hy ? ctx.v1 : ctx.v};
            if (n > 1) hipLaunchKernelGGL(screen_saved, dim3((n - 1 + 255) / 256), dim3(256), 0, ctx.stream, sa);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipMemcpyAsync(count, ctx.devcount, sizeof(int), hipMemcpyDeviceToHost, ctx.stream));
            HIPCHECK(hipStreamSynchronize(ctx.stream));
            if (*count <= ctx.ev_cap) break;
            ensure_events(*count);   // screen again after growth: nothing is dropped
        }
        fetch_events(*count);
        std::sort(ctx.hev.begin(), ctx.hev.end(), [](const DevEvent& a, const DevEvent& b) {
            return pair_less(a.i, a.j, b.i, b.j);
        });
        ctx.events.resize(*count);
        for (int k = 0; k < *count; k++) ctx.events[k] = to_public(ctx.hev[k]);
        *out = ctx.events.data();
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int mercury_hip_force(double* a) {
    try {
        require_state();
        radau_flush();
        if (ctx.systems.size() != 1) throw std::runtime_error("force needs one system");
        ensure_stream();
        clear_fault();
        if (ensure_radau() && ctx.algorithm == 4) ctx.history = false;
        build_systems(kRaEval, true, 1.0, 0.0);
        radau_pass(0.0, 0, false);
        check_fault("Nonfinite force");
        download_vec3(ctx.ra1, a);
        a[0] = a[1] = a[2] = 0.0;
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int mercury_hip_force_noalloc(int n, int nbig, int pn, const double* m, const double* x, const double* v,
                              const double* jcen, double* a, long long ld) {
    try {
        if (n < 1 || nbig < 1 || nbig > n) throw std::runtime_error("invalid body counts");
        if (nbig > kMaxBig) throw std::runtime_error("more Big bodies than the allocation-free force's scratch holds (64)");
        if (ld < n) throw std::runtime_error("leading dimension smaller than the body count");
        if (!m || !x || !a || (pn && !v) || !jcen) throw std::runtime_error("missing array");
        int dev = 0;
        HIPCHECK(hipGetDevice(&dev));
        if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index beyond the scratch table");
        if (!scratch_ptr[dev])
            HIPCHECK(hipGetSymbolAddress(reinterpret_cast<void**>(&scratch_ptr[dev]), HIP_SYMBOL(scratch)));
        unsigned char* scr = scratch_ptr[dev];
        double mu = 0.0;
        HIPCHECK(hipMemcpy(&mu, m, sizeof(double), hipMemcpyDeviceToHost));
        if (!(mu > 0)) throw std::runtime_error("central mass must be positive");
        const int np = nbig - 1;
        Sys d{};
        d.base = 0; d.n = n; d.nbig = nbig; d.ts = start_stride(np);
        d.tab = 0; d.start = 0;
        d.mu = mu; d.h = 1.0; d.hby2 = 0.5; d.rcen = 0.0;
        d.j2 = jcen[0]; d.j4 = jcen[1]; d.j6 = jcen[2];
        d.obl = (jcen[0] != 0 || jcen[1] != 0 || jcen[2] != 0);
        State st{};
        st.x = const_cast<double*>(x); st.v = const_cast<double*>(pn ? v : x); st.m = m;
        st.fault = reinterpret_cast<int*>(scr + kNoAllocFault);
        st.cap = ld;
        static const RaConst K = ra_constants();
        RaArgs args{reinterpret_cast<const Sys*>(scr), st, RaState{nullptr, nullptr, a},
                    reinterpret_cast<double*>(scr + kNoAllocTab), nullptr, K, 0.0, 0, 0, pn ? 1 : 0,
                    light_speed2()};
        hipLaunchKernelGGL(noalloc_setup, dim3(1), dim3(64), 0, 0, d, scr, a, ld);
        if (pn) hipLaunchKernelGGL((ra_planets<true, false>), dim3(1), dim3(64), 0, 0, args);
        else hipLaunchKernelGGL((ra_planets<false, false>), dim3(1), dim3(64), 0, 0, args);
        if (n > nbig) {
            dim3 grid((n - nbig + 127) / 128);
            if (pn) hipLaunchKernelGGL((ra_particles<true, false>), grid, dim3(128), 0, 0, args);
            else hipLaunchKernelGGL((ra_particles<false, false>), grid, dim3(128), 0, 0, args);
        }
        HIPCHECK(hipGetLastError());
        int fault = 0;
        HIPCHECK(hipMemcpy(&fault, scr + kNoAllocFault, sizeof(int), hipMemcpyDeviceToHost));
        if (fault & FAULT_FORCE) throw std::runtime_error("Nonfinite force: non-finite acceleration");
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int mercury_hip_export(double h, int physical, double* x, double* v) {
    try {
        require_state();
        radau_flush();
        if (physical && ctx.algorithm == 1) {
            // mco_mvs2h on scratch (mercury6_2.for:2107-2219); the state is untouched
            if (ctx.systems.size() != 1) throw std::runtime_error("corrector needs one system");
            ensure_stream();
            build_systems(std::max(ctx.steps_per_launch, 4), true, h, 0.0);
            if (!ctx.cx) { dalloc(ctx.cx, 3 * ctx.cap); dalloc(ctx.cv, 3 * ctx.cap); }
            double rt10 = std::sqrt(10.0);
            CorrectorArgs ca{ctx.dsys, device_state(), ctx.tab, ctx.cx, ctx.cv,
                             {h * rt10 * 3.0 / 10.0, h * rt10 / 5.0}, {-h * rt10 / 72.0, h * rt10 / 24.0}};
            hipLaunchKernelGGL(corrector_planets, dim3(1), dim3(64), 0, ctx.stream, ca);
            int npart = ctx.n - ctx.nbig;
            if (npart > 0 && ctx.comets.empty())
                hipLaunchKernelGGL(corrector_particles<false>, dim3((npart + 127) / 128), dim3(128), 0, ctx.stream, ca);
            else if (npart > 0)
                hipLaunchKernelGGL(corrector_particles<true>, dim3((npart + 127) / 128), dim3(128), 0, ctx.stream, ca);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipStreamSynchronize(ctx.stream));
            download_vec3(ctx.cx, x);
            download_vec3(ctx.cv, v);
            return 0;
        }
        download_vec3(ctx.x, x);
        download_vec3(ctx.v, v);
        if (physical && ctx.algorithm == 10) {
            // mco_dh2h (mercury6_2.for:2007-2063): heliocentric velocities
            std::vector<double> m((size_t)ctx.n);
            HIPCHECK(hipMemcpy(m.data(), ctx.m, sizeof(double) * ctx.n, hipMemcpyDeviceToHost));
            double mv[3] = {0.0, 0.0, 0.0};
            for (int j = 1; j < ctx.n; j++)
                for (int k = 0; k < 3; k++) mv[k] = mv[k] + m[j] * v[3 * j + k];
            double temp = 1.0 / m[0];
            for (int k = 0; k < 3; k++) mv[k] = temp * mv[k];
            for (int j = 1; j < ctx.n; j++)
                for (int k = 0; k < 3; k++) v[3 * j + k] = v[3 * j + k] + mv[k];
        }
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int mercury_hip_hybrid_begin(double h, const double* crit, int flag, int cap, int* ce, int* nce, int* pi, int* pj,
                             double* x, double* v, int64_t* forces) {
    try {
        require_state();
        hybrid_begin(h, crit, flag, cap, ce, nce, pi, pj, x, v, forces);
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int mercury_hip_hybrid_finish(double h, const double* m, const double* x, const double* v, int64_t* forces) {
    try {
        require_state();
        hybrid_finish(h, m, x, v, forces);
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int mercury_hip_encounter_enter(int n, int nbig, const double* m, const double* x, const double* v,
                                const double* crit, const double* rce, const double* rphys, int np, const int* pi,
                                const int* pj) {
    try {
        (void)rce;
        (void)rphys;
        encounter_enter(n, nbig, m, x, v, crit, np, pi, pj);
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int mercury_hip_encounter_update(const double* m, const double* x, const double* v) {
    try {
        Context::Enc& e = ctx.enc;
        if (!e.active) throw std::runtime_error("no encounter context");
        enc_settle();
        for (int j = e.nbig; j < e.n; j++)
            if (m[j] != 0) throw std::runtime_error("Small bodies with mass are not provided by the HIP path");
        HIPCHECK(hipMemcpy(e.m, m, sizeof(double) * e.n, hipMemcpyHostToDevice));
        encounter_pack();
        encounter_states(x, v);
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

void mercury_hip_encounter_exit(void) {
    if (ctx.enc.pending) (void)hipStreamSynchronize(ctx.stream);   // no error crosses this void entry point
    ctx.enc.pending = false;
    ctx.enc.active = false;
}

int mercury_hip_encounter_advance(double h0, double* hrec, double tol, int64_t* substeps, int64_t* forces,
                                  int64_t* rejected) {
    try {
        encounter_advance(h0, hrec, tol, substeps, forces, rejected);
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

void mercury_hip_free(void) { release_all(); }

// ---- extensions -----------------------------------------------------------

int mercury_hip_advance(double time, double h, int nsteps, double rcen, int screen, int64_t* forces) {
    try {
        require_state();
        if (!mvs_family()) throw std::runtime_error("advance needs MVS");
        if (nsteps < 0) throw std::runtime_error("negative step count");
        if (!ctx.ensemble && (!(h != 0) || !std::isfinite(h))) throw std::runtime_error("invalid step");
        ensure_stream();
        ctx.tape.valid = false;
        ctx.adv_events.clear();
        bool reset = !ctx.history;
        int cnt = 0;
        if (nsteps > 0) {
            // the call's first kernel clears the fault flag and the event count;
            // its last writes them, with the slow-set size, to ctx.hstatus
            mvs_run(time, h, nsteps, reset, screen != 0, false, rcen);
            ctx.history = true;
            const int f = ctx.hstatus[0];
            check_ngf_fault(f);
            if (f & FAULT_FORCE) throw std::runtime_error("MVS force failed: non-finite acceleration");
            if (f & FAULT_REQ_OVERFLOW)
                throw std::runtime_error("screening request buffer overflow: use fewer steps per launch");
            if (f & FAULT_EVENT_OVERFLOW) {
                ensure_events(ctx.hstatus[1]);   // room for as many in the next call
                throw std::runtime_error("event buffer overflow: advance fewer steps per call");
            }
            cnt = ctx.hstatus[1];
            ctx.parked = ctx.hstatus[2];
        }
        fetch_events(std::min(cnt, ctx.ev_cap));
        for (auto& e : ctx.hev) {
            MercuryHipEvent r;
            r.e = to_public(e);
            r.step = e.step;
            r.system = e.sys;
            ctx.adv_events.push_back(r);
        }
        ctx.saved_valid = false;
        std::sort(ctx.adv_events.begin(), ctx.adv_events.end(), [](const MercuryHipEvent& a, const MercuryHipEvent& b) {
            if (a.system != b.system) return a.system < b.system;
            if (a.step != b.step) return a.step < b.step;
            return pair_less(a.e.i, a.e.j, b.e.i, b.e.j);
        });
        *forces = nsteps + ((reset && nsteps > 0) ? 1 : 0);
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int mercury_hip_ensemble_upload(int nsys, const int* n, const int* nbig, const double* m, const double* x,
                                const double* v, const double* rce, const double* rphys, const double* h,
                                const double* rcen) {
    try {
        if (nsys < 1) throw std::runtime_error("no systems");
        std::vector<Context::HostSys> sys(nsys);
        long long total = 0;
        for (int s = 0; s < nsys; s++) {
            if (n[s] < 1 || nbig[s] < 1 || nbig[s] > n[s]) throw std::runtime_error("invalid body counts");
            if (nbig[s] > kMaxBig) throw std::runtime_error("more Big bodies than the HIP path supports (64)");
            if (!(m[total] > 0)) throw std::runtime_error("central mass must be positive");
            if (!(h[s] != 0) || !std::isfinite(h[s])) throw std::runtime_error("invalid step");
            sys[s] = Context::HostSys{(int)total, n[s], nbig[s], m[total], h[s], rcen[s], 0.0, 0.0, 0.0};
            total += n[s];
        }
        if (total > (1LL << 30)) throw std::runtime_error("too many bodies");
        for (long long j = 0; j < 3 * total; j++)
            if (!std::isfinite(x[j]) || !std::isfinite(v[j])) throw std::runtime_error("nonfinite state");
        if (ctx.algorithm != 1 && ctx.algorithm != 9) mercury_hip_configure(1);
        ensure_stream();
        ensure_capacity((int)total);
        ctx.n = (int)total; ctx.nbig = 0; ctx.pn = 0; ctx.ngflag = 0;
        ctx.comets.clear();
        for (int k = 0; k < 3; k++) ctx.jcen[k] = 0.0;
        ctx.systems = sys;
        ctx.ensemble = true;
        ctx.sys_valid = false;
        ctx.upload_gen++;
        clear_slow_set();
        HIPCHECK(hipMemcpy(ctx.m, m, sizeof(double) * total, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(ctx.rce, rce, sizeof(double) * total, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(ctx.rphys, rphys, sizeof(double) * total, hipMemcpyHostToDevice));
        upload_vec3(x, ctx.x);
        upload_vec3(v, ctx.v);
        HIPCHECK(hipMemset(ctx.a, 0, sizeof(double) * 3 * ctx.cap));
        HIPCHECK(hipMemset(ctx.ox, 0, sizeof(double) * 3 * ctx.cap));
        HIPCHECK(hipMemset(ctx.ov, 0, sizeof(double) * 3 * ctx.cap));
        HIPCHECK(hipMemset(ctx.fault, 0, sizeof(int)));
        build_tiles();
        ctx.history = false;
        ctx.saved_valid = false;
        ctx.have_state = true;
        ctx.tape.valid = false;
        HIPCHECK(hipDeviceSynchronize());
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int mercury_hip_ensemble_advance(double time, int nsteps, int screen, int64_t* forces) {
    return mercury_hip_advance(time, 0.0, nsteps, 0.0, screen, forces);
}

int mercury_hip_ensemble_download(double* x, double* v) { return mercury_hip_download(x, v, 0); }

int mercury_hip_advance_events(const MercuryHipEvent** events, int* count) {
    *events = ctx.adv_events.data();
    *count = (int)ctx.adv_events.size();
    return 0;
}

int mercury_hip_set_tuning(int steps_per_launch, int block_size, int particles_per_lane) {
    try {
        if (steps_per_launch < 0 || block_size < 0 || particles_per_lane < 0) throw std::runtime_error("invalid tuning");
        if (block_size && (block_size % 64 || block_size > 256))
            throw std::runtime_error("block size must be 64..256, multiple of 64");
        if (particles_per_lane > kMaxPpl) throw std::runtime_error("particles per lane must be 1..8");
        if (steps_per_launch) ctx.steps_per_launch = steps_per_launch;
        if (block_size) ctx.block = block_size;
        if (particles_per_lane) ctx.ppl = particles_per_lane;
        if ((block_size || particles_per_lane) && ctx.have_state) build_tiles();
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int mercury_hip_timing(int enable) {
    try {
        ctx.timing.clear();
        ctx.timing.on = enable != 0;
        ctx.timing.fast_only = enable == 2;
        ctx.work_total = 0;
        if (ctx.timing.on) {
            if (!ctx.work) dalloc(ctx.work, 2);
            HIPCHECK(hipMemset(ctx.work, 0, 2 * sizeof(unsigned long long)));
        }
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int mercury_hip_timing_read(double* particle_ms, int* particle_launches, double* planet_ms, int* planet_launches,
                            double* general_ms, int* general_launches) {
    try {
        if (ctx.stream) {
            HIPCHECK(hipStreamSynchronize(ctx.stream));
            HIPCHECK(hipStreamSynchronize(ctx.sm));
        }
        double ms[3] = {0, 0, 0};
        int cnt[3] = {0, 0, 0};
        for (auto& mk : ctx.timing.marks) {
            float t = 0;
            HIPCHECK(hipEventElapsedTime(&t, mk.second.first, mk.second.second));
            ms[mk.first] += t;
            cnt[mk.first]++;
        }
        *particle_ms = ms[K_PARTICLE]; *particle_launches = cnt[K_PARTICLE];
        *planet_ms = ms[K_PLANET]; *planet_launches = cnt[K_PLANET];
        *general_ms = ms[K_GENERAL]; *general_launches = cnt[K_GENERAL];
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int64_t mercury_hip_parked(void) { return ctx.parked; }

int mercury_hip_adjoint_forward(int nsteps, double h, int segment, double* x, double* v) {
    try {
        require_state();
        clear_fault();
        adjoint_forward(nsteps, h, segment, x, v);
        check_fault("MVS force failed");
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int mercury_hip_adjoint_backward(const double* gx, const double* gv, double* gx0, double* gv0, double* gm) {
    try {
        require_state();
        adjoint_backward(gx, gv, gx0, gv0, gm);
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int mercury_hip_tangent(const double* dx0, const double* dv0, const double* dm, double* dx, double* dv) {
    try {
        require_state();
        tangent_pass(dx0, dv0, dm, dx, dv);
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int mercury_hip_set_graph(int enable) {
    graphs.enabled = enable != 0;
    if (!graphs.enabled) graphs.drop();
    return 0;
}

int mercury_hip_graph_stats(int* captures, int* replays) {
    *captures = graphs.captures;
    *replays = graphs.replays;
    return 0;
}

int mercury_hip_autotune(double h, int probe_steps, int* spl_out, int* block_out, int* ppl_out, double* rate_out) {
    Context& c = ctx;
    double* save = nullptr;
    double* tsave = nullptr;
    try {
        require_state();
        if (!mvs_family()) throw std::runtime_error("autotuning needs the MVS algorithm");
        if (probe_steps < 1) throw std::runtime_error("probe needs at least one step");
        const bool single = !c.ensemble;
        if (single && (!(h != 0) || !std::isfinite(h))) throw std::runtime_error("invalid step");
        ensure_stream();
        const int nsys = (int)c.systems.size();
        ensure_systems(nsys);
        // the state is put back after every probe
        dalloc(save, (size_t)9 * c.cap);
        dalloc(tsave, (size_t)nsys);
        double* arr[3] = {c.x, c.v, c.a};
        for (int q = 0; q < 3; q++)
            HIPCHECK(hipMemcpy(save + (size_t)q * 3 * c.cap, arr[q], sizeof(double) * 3 * c.cap, hipMemcpyDeviceToDevice));
        HIPCHECK(hipMemcpy(tsave, c.tcur, sizeof(double) * nsys, hipMemcpyDeviceToDevice));
        const bool hist = c.history;
        const int spl0 = c.steps_per_launch, block0 = c.block, ppl0 = c.ppl;
        auto restore = [&]() {
            for (int q = 0; q < 3; q++)
                HIPCHECK(hipMemcpy(arr[q], save + (size_t)q * 3 * c.cap, sizeof(double) * 3 * c.cap,
                                   hipMemcpyDeviceToDevice));
            HIPCHECK(hipMemcpy(c.tcur, tsave, sizeof(double) * nsys, hipMemcpyDeviceToDevice));
            c.history = hist;
            clear_slow_set();   // the probe's slow set belongs to the probe's end state
        };
        const int spls[4] = {8, 16, 32, 64};
        const int blocks[3] = {64, 128, 256};
        const int ppls[3] = {1, 2, 4};
        double best = -1.0;
        int bspl = spl0, bblock = block0, bppl = ppl0;
        const double rc = single ? c.systems[0].rcen : 0.0;
        for (int pi = 0; pi < 3; pi++)
          for (int bi = 0; bi < 3; bi++)
            for (int si = 0; si < 4; si++) {
                c.steps_per_launch = spls[si];
                c.block = blocks[bi];
                c.ppl = ppls[pi];
                build_tiles();
                double sec = 0.0;
                for (int rep = 0; rep < 2; rep++) {   // the first run warms up (and captures a graph)
                    restore();
                    clear_fault();
                    HIPCHECK(hipMemset(c.devcount, 0, sizeof(int)));
                    HIPCHECK(hipDeviceSynchronize());
                    auto t0 = std::chrono::steady_clock::now();
                    mvs_run(0.0, h, probe_steps, !c.history, true, false, rc);
                    sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                }
                check_fault("MVS force failed");
                double rate = (double)c.npart * probe_steps / sec;
                if (rate > best) { best = rate; bspl = spls[si]; bblock = blocks[bi]; bppl = ppls[pi]; }
            }
        c.steps_per_launch = bspl;
        c.block = bblock;
        c.ppl = bppl;
        build_tiles();
        restore();
        HIPCHECK(hipDeviceSynchronize());
        dfree(save);
        dfree(tsave);
        *spl_out = bspl;
        *block_out = bblock;
        if (ppl_out) *ppl_out = bppl;
        *rate_out = best;
        return 0;
    } catch (const std::exception& e) {
        dfree(save);
        dfree(tsave);
        return fail(e);
    }
}

int mercury_hip_check_arith(int64_t n, uint64_t seed, int64_t* mismatches) {
    unsigned long long* bad = nullptr;
    try {
        if (n < 0) throw std::runtime_error("negative sample count");
        dalloc(bad, 4);
        HIPCHECK(hipMemset(bad, 0, 4 * sizeof(unsigned long long)));
        const long long blocks = std::min<long long>(std::max<long long>((n + 255) / 256, 1), 8192);
        hipLaunchKernelGGL(check_arith, dim3((unsigned)blocks), dim3(256), 0, 0, (long long)n,
                           (unsigned long long)seed, bad);
        HIPCHECK(hipGetLastError());
        unsigned long long h[4];
        HIPCHECK(hipMemcpy(h, bad, sizeof(h), hipMemcpyDeviceToHost));
        dfree(bad);
        for (int q = 0; q < 4; q++) mismatches[q] = (int64_t)h[q];
        return 0;
    } catch (const std::exception& e) {
        dfree(bad);
        return fail(e);
    }
}

int mercury_hip_work_read(int64_t* fast_steps, int64_t* general_steps) {
    try {
        unsigned long long w[2] = {0, 0};
        if (ctx.work) HIPCHECK(hipMemcpy(w, ctx.work, sizeof(w), hipMemcpyDeviceToHost));
        *fast_steps = (int64_t)ctx.work_total - (int64_t)w[1];
        *general_steps = (int64_t)w[1];
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

}  // extern "C"
