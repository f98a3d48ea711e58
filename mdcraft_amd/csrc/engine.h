// Host-side engine state shared by the kernel launchers and the C ABI.
#pragma once
#include <string>
#include <tuple>
#include <vector>
#include <stdexcept>
#include <hip/hip_ext.h>
#include "common.h"

namespace mdc {

enum ErrCode { E_OK = 0, E_INVAL = 1, E_HIP = 2, E_NOMEM = 3, E_FILE = 4, E_STATE = 5, E_BUDGET = 6, E_RANGE = 7 };

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHECK(x)                                                                                  \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) throw ::mdc::Error(::mdc::E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Per-atom double fields, SoA with stride N (dstate[f*N + i]).
enum DField {
    D_VX, D_VY, D_VZ,        // velocity (holds the half-step velocity inside a step sequence)
    D_FX, D_FY, D_FZ,        // FP64 force
    D_M, D_RNS,              // mass, neighbour search radius
    D_EP, D_EM, D_NC, D_FP,  // pair energy, embedding energy, EAM density, F'(density)
    D_VXX, D_VYY, D_VZZ, D_VXY, D_VXZ, D_VYZ,  // per-atom virial (symmetric part)
    D_NFIELDS
};
// Per-atom float fields (mixed-precision force)
enum FField { F_FX, F_FY, F_FZ, F_NFIELDS };

enum Prec { P_FP64 = 0, P_MIXED = 1 };
// W_NC_STORED: the EAM density pass takes F(nc), F'(nc) of the stored
// density instead of the one it sums, and leaves nc as it is
enum What { W_FORCE = 1, W_ENERGY = 2, W_VIRIAL = 4, W_NC_STORED = 8 };
// solver calls of the reference's Python API, one at a time
enum SolverStage { ST_PREPARE = 0, ST_CHARGE = 1, ST_FORCES = 2, ST_VIRIALS = 3 };
enum PotType { PT_LJ = 1, PT_EAM = 2, PT_BPNN = 4 };
// misc device counters
// M_RMIN: the least squared pair distance (FP32 bits) the fixed-point LDS
// passes have met since positions were last uploaded; it sets the part of
// the table they stage
enum Misc { M_MAXNN = 0, M_ABORT = 1, M_ERR = 2, M_DECIDE = 3, M_RMIN = 4, M_COUNT = 8 };
constexpr int kRminNone = 0x7f800000;  // +inf: nothing met yet
enum ErrBits { EB_RHO_RANGE = 1, EB_NONFINITE = 2, EB_R_RANGE = 4 };

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // pinned host mirrors of the misc counters: slot 0 for synchronous
    // reads, slots 1 and 2 for the two step chunks kept in flight, each
    // with the event that marks its copy done
    int* h_pin = nullptr;
    hipEvent_t chunk_ev[2] = {nullptr, nullptr};
    int cu_count = 0;
    size_t lds_max = 65536;  // LDS bytes one workgroup may use
    // waits on the stream poll it (a host core kept busy) instead of
    // blocking: a shorter wake-up for short evaluations
    bool spin = false;
};

// Waits for the context's stream, or for an event on it, by blocking or,
// with Ctx::spin, by polling (forces.hip).
void stream_wait(Ctx& c);
void event_wait(Ctx& c, hipEvent_t e);

// Langevin thermostats (thermostat/langevin.cxx:7-42) applied one after
// the other, as thermostat::Compound does
constexpr int kMaxTherm = 4;
struct LangevinP {
    double beta, sigma, xmin, xmax;
    double U[3], heat[3];
};
struct ThermArgs {
    int count;
    int pad;
    uint64_t seed;
    LangevinP t[kMaxTherm];
};

// Pair potentials: fn 0 = LJsCut (LJs-cut.cxx:8-20), 1 = LJs, the LJ with
// a polynomial tail that takes V and V' to zero at rc (LJs.cxx:7-26,
// coefficients as its constructor leaves them), 2 = tabulated Pair
// (potential/pair.cxx; spline records in the Pot's EamTables r64/r32)
struct LJParams {
    int fn = 0;
    double a, s, s6, lj1, lj2, factor, rc, rc2;
    double aVr, aVr2, rVr1, rVr2, aLJ, aLJ3, bLJ, bLJ2;
};

// One EAM object of a Multi potential, used for one pair of kinds
// (solver::Multi, multi.cxx:26-49): offsets of its rho and r*phi interval
// records in the r-record array and of its F records, and its grids.
struct EamObj {
    // pair_only: a tabulated Pair member (no density, no embedding; its
    // virial leaves Em, pair.cxx:235-324).
    // rho_off / F_off: the density table that an atom of the lower kind of
    // the pair sums over its neighbour and that atom's embedding table;
    // rho_hi_off / F_hi_off: the same for the atom of the higher kind.  A
    // Multi object uses one density and one embedding table for both atoms
    // (the offsets agree); an alloy object of elements a < b holds rho_b and
    // F_a, then rho_a and F_b.
    int rho_off, phi_off, F_off, nr, nrho, pair_only, rho_hi_off, F_hi_off;
    double dr, dr1, drho, drho1, rc2;
};
constexpr int kMaxMultiKinds = 4;
constexpr int kMaxPairs = kMaxMultiKinds * (kMaxMultiKinds + 1) / 2;
__host__ __device__ inline int pair_index(int a, int b) {
    const int hi = a > b ? a : b, lo = a > b ? b : a;
    return hi * (hi + 1) / 2 + lo;
}

// Spline tables of an EAM potential: one element, or a Multi potential of
// one object per pair of kinds.  Interval record k of a table holds
// (y[k], z[k], y[k+1], z[k+1]); the record past the end repeats the last
// knot, as the reference's [N1+1] read at N1 = N-1 assumes a value.
struct EamTables {
    int nk = 1;
    int nr = 0, nrho = 0;
    double dr = 0, drho = 0, rc = 0;
    // device: per object the rho records [nr] then the r*phi records [nr];
    // embedding records [nrho] per object
    double4* r64 = nullptr;
    float4* r32 = nullptr;
    double4* F64 = nullptr;
    // one element: the FP32 rho and r*phi records of each interval side by
    // side ([2 nr]: rho k, r*phi k), so the force pass reads both in one
    // 32-byte record
    float4* r32p = nullptr;
    // Multi potential (nk > 1): objects per pair of kinds, records of all
    // objects concatenated in r64 / r32 / F64.  alloy: the nk elements of an
    // alloy EAM (per-element density and embedding, a pair term per pair of
    // elements) in the same objects, whose two sides then differ.
    int npair = 0;
    bool alloy = false;
    size_t nrec = 0, nFrec = 0;
    EamObj* objs = nullptr;
};

struct BpnnParams;

struct Pot {
    Ctx* ctx = nullptr;
    int type = 0;
    double rcut = 0;
    LJParams lj{};
    EamTables eam;
    BpnnParams* bp = nullptr;
};

struct Batch {
    Ctx* ctx = nullptr;
    int nsys = 0;
    int n = 0;
    std::vector<SysDesc> desc;
    SysDesc* d_desc = nullptr;
    std::vector<int> atom_off;  // [nsys+1]
    // records of each box that hold atoms; the rest of the box's records
    // are void (mdc_sys_refill_box)
    std::vector<int> present;
    // Sorted order: every box's live records in cell order, box after box,
    // in [0, live), and the void records of all boxes after them (one void
    // cell past every box's grid).  Per-atom kernels run over [0, live);
    // n stays the stride of the per-field arrays.  Both position buffers
    // and both copies of the field arrays hold the same void tail, so the
    // stepping kernels never touch it.  full_sort: records became void or
    // live since the last sort (refill, upload over voids), so the next
    // rebuild sorts all n records and copies the new tail; otherwise it
    // sorts [0, live) only.
    int live = 0;
    bool full_sort = true;
    // state
    double4* pos[2] = {nullptr, nullptr};
    int cur = 0;
    double* dst[2] = {nullptr, nullptr};
    float* fst[2] = {nullptr, nullptr};
    int* orig[2] = {nullptr, nullptr};
    int sc = 0;  // which dst/fst/orig copy is current
    double4* rsave = nullptr;
    // single-element EAM: each atom's position with F'(nc) in w, written by
    // the density pass so the force pass gathers both in one record
    double4* posF = nullptr;
    // Fixed-point step path (single-element EAM in mixed precision, every
    // box with a unique minimum image): uq[c] holds the fixed-point record
    // (SysDesc qs) of pos[c], written with every drift and after every
    // sort; uqF the same coordinates with F'(nc) as FP32 in w, written by
    // the density pass.  Allocated on first use (uq_n records).
    uint4* uq[2] = {nullptr, nullptr};
    uint4* uqF = nullptr;
    int uq_n = 0;
    bool qfast = false;    // set while a step sequence or a kernel timing runs on this path
    bool q_off = false;    // mdc_sys_step_path(1): never take it
    bool q_lds = true;     // the variants that stage the table in LDS
    int q_last = 0;        // the last step sequence or timing: 0 not taken, 1 global tables, 2 tables in LDS
    bool all_shift = false;  // every box's shift_ok (set_trigger)
    int* where = nullptr;  // original index -> sorted index
    // cells
    int ncells = 0;
    int cell_alloc = 0;  // cells the cell arrays hold (grown, never shrunk)
    int* cell_count = nullptr;
    int* cell_start = nullptr;
    int* cell_fill = nullptr;
    int* atom_cell = nullptr;
    int* sorted_src = nullptr;
    // list
    int tpa = 4;
    int cap = 0;   // slots per atom (multiple of tpa)
    int M = 0;     // cap / tpa
    size_t nl_elems = 0;
    int* nl = nullptr;
    int* nn = nullptr;
    int* misc = nullptr;
    bool list_valid = false;
    // positions were uploaded after the last build: the image shifts stored
    // in the list entries may not match them until the next build
    bool shifts_stale = false;
    int f_prec = -1;      // which force array holds the current force
    double last_rcut = -1;
    int rebuilds = 0;
    // list rebuilt only on request (the reference's VerletList::update);
    // step sequences then never trigger one and choose images per step
    bool manual_rebuild = false;
    // precision of the BPNN descriptors and dE/dG of the last prepare stage
    // attached thermostats and the number of applications so far (the
    // counter of the random stream)
    ThermArgs therm{};
    uint64_t therm_calls = 0;
    int bp_prepared = -1;
    int bp_prepared_builds = -1;
    // steps between rebuilds (running estimate) and steps since the last one,
    // used to size the launch chunks of a step sequence
    double pred_interval = 0.0;
    int since_rebuild = 0;
    // host copies of the integer fields the device does not need
    std::vector<uint16_t> kind, tag;
    // largest atom kind held (-1: to be recounted after kinds were written)
    int kind_max = -1;
    std::vector<uint32_t> uid;
    std::vector<double> rcut_host;
    std::vector<uint8_t> periodic_bc;  // per system: solver built with the Periodic boundary
    // F'(nc_i) of every atom under each object of a Multi potential: [nk][n]
    double* fpm = nullptr;
    size_t fpm_elems = 0;
    // scratch for BPNN
    void* bp_scratch = nullptr;
    size_t bp_scratch_bytes = 0;
    // set for the duration of an allocation-free evaluation: the BPNN work
    // arrays are carved from this device work area instead
    char* bp_fixed = nullptr;
    // reduction output
    double* thermo = nullptr;
    // device staging of record transfers (upload, refill, download), grown
    // as needed and kept, so transfers allocate and free nothing
    char* stage = nullptr;
    size_t stage_bytes = 0;
    // Reallocations of any device array the evaluation kernels take (list,
    // cells, work arrays): a captured evaluation is valid only for the
    // arrays it was captured with.
    uint64_t alloc_gen = 0;
    // Graph mode of mdc_compute: the device work of an evaluation captured
    // once per configuration (key) and replayed.  A configuration's first
    // call runs as usual (so that every allocation happens outside a
    // capture), the second captures it, later ones replay the graph.
    bool graphs = false;
    struct GraphKey {
        const void* pot = nullptr;
        int what = 0, prec = 0, rebuilds = 0, cap = 0, live = 0, tpa = 0, block = 0, cur = 0, sc = 0;
        uint64_t alloc_gen = 0;
        bool operator==(const GraphKey& o) const {
            return pot == o.pot && what == o.what && prec == o.prec && rebuilds == o.rebuilds && cap == o.cap &&
                   live == o.live && tpa == o.tpa && block == o.block && cur == o.cur && sc == o.sc &&
                   alloc_gen == o.alloc_gen;
        }
    };
    GraphKey gkey;        // configuration of gexec, or of the last eager call
    bool gkey_set = false;
    hipGraphExec_t gexec = nullptr;
    long graph_launches = 0, graph_captures = 0;
    // threads per workgroup of the pair-potential kernels (64, 128 or 256)
    int block = 256;
    // per-system extent of rns (host), for the displacement trigger
    std::vector<double> rns_min, rns_max;
    // optional kernel timing: hipEvents recorded on the launch stream around
    // each launch of the timed kernels, drained at the next sync point
    bool timing = false;
    std::vector<hipEvent_t> ev_free;
    struct Timed { hipEvent_t a, b; int kind; };
    std::vector<Timed> ev_pending;
    double kt_ms[9] = {};  // per timed kernel kind (KKind)
    long kt_n[9] = {};

    double* d(int s, int f) const { return dst[s] + (size_t)f * n; }
    float* fl(int s, int f) const { return fst[s] + (size_t)f * n; }
};

// timed kernel kinds
enum KKind {
    K_FORCE = 0, K_DENSITY = 1, K_PREDRIFT = 2, K_BP_DESC = 3, K_BP_MLP = 4, K_BP_FORCE = 5, K_BP_GATHER = 6, K_REBUILD = 7,
    K_BP_FUSED = 8
};
constexpr int kKinds = 9;
static_assert(sizeof(Batch::kt_n) / sizeof(long) == kKinds, "one timing slot per kernel kind");
hipEvent_t timer_begin(Batch& b);
void timer_end(Batch& b, hipEvent_t a, int kind);
void timer_drain(Batch& b);
// drops the batch's captured evaluation (forces.hip)
void graph_release(Batch& b);
// species of a BPNN potential's model (bpnn.hip)
int bpnn_species(const Pot& p);
hipEvent_t timer_event(Batch& b);

// Launches kernel k on the context's stream.  While timing is on, the
// start/stop events travel inside the kernel's own dispatch
// (hipExtLaunchKernel), so the recorded time is the kernel's execution
// without the cost of separate marker packets.
template <typename... P, typename... A>
void launch_k(Batch& b, int kind, void (*k)(P...), dim3 grid, dim3 block, unsigned lds, A... args) {
    static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
    std::tuple<P...> t{static_cast<P>(args)...};
    void* argv[sizeof...(P) > 0 ? sizeof...(P) : 1];
    std::apply([&](auto&... x) {
        int i = 0;
        ((argv[i++] = (void*)&x), ...);
    }, t);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (b.timing && kind >= 0) {
        e0 = timer_event(b);
        e1 = timer_event(b);
    }
    const hipError_t err = e0 ? hipExtLaunchKernel((const void*)k, grid, block, argv, lds, b.ctx->stream, e0, e1, 0)
                              : hipLaunchKernel((const void*)k, grid, block, argv, lds, b.ctx->stream);
    if (err != hipSuccess) throw Error(E_HIP, std::string("kernel launch: ") + hipGetErrorString(err));
    if (e0) b.ev_pending.push_back({e0, e1, kind});
}

// launchers (nlist.hip)
void nlist_rebuild(Batch& b, bool sync = true);
void nlist_regrow(Batch& b, int maxnn);
int nlist_decide(Batch& b, double rcut);
// launchers (forces.hip)
void force_compute(Batch& b, const Pot& p, int what, int prec);
void run_steps(Batch& b, const Pot& p, double dt, int nsteps, int prec, int chunk);
void time_step_kernels(Batch& b, const Pot& p, int prec, double dt, int reps, double* us);
void thermo(Batch& b, double* out);
void set_trigger(Batch& b, double rcut, bool allow_shift);
void stage_compute(Batch& b, const Pot& p, int stage, int prec);
void check_errors(Batch& b, const int* counters = nullptr);
// measurement hook (calib.hip)
void copy_calibration(int device, size_t bytes, int reps, double* us_per_copy);
// launchers (bpnn.hip)
// parts: 1 = descriptors and network (G, dE/dG, Ep), 2 = forces from the
// stored dE/dG, 3 = both
void bpnn_compute(Batch& b, const Pot& p, int what, int prec, int mode, double dt, int step, const double4* pos,
                  double4* pos_next, int parts = 3);
void bpnn_release_work(Batch& b);
void bpnn_inspect(Batch& b, const Pot& p, double* G, double* dEdG);
int bpnn_inputs(const Pot& p);
// bytes of BPNN work arrays one evaluation of this batch needs when they are
// carved from one work area (each array 256-byte aligned)
size_t bpnn_fixed_bytes(const Batch& b, const Pot& p, int what, int prec);
Pot* bpnn_create(Ctx* ctx, const char* path);
// architecture (nin, h1, h2, 1), descriptor counts (radial, angular, G4 eta
// values, (lambda, zeta) groups) and cutoff of a MAISE file, host only
void bpnn_describe(const char* path, int* arch, int* counts, double* rcut);
void bpnn_destroy(Pot* p);

}  // namespace mdc
