// Behler-Parrinello network (MAISE model format) on the device.
//
// Reference semantics (maise_bpnn.cxx:122-325, tools/bpnn/base.h:95-190,
// base.cxx:116-128, maise_ann.cxx:8-55, parser/maise_model.cxx:29-129):
//   G2_r(i)   = sum_j exp(-eta_r (r_ij - Rs)^2) fc(r_ij)
//   G4_a(i)   = sum_{j<k} 2 * 2^(1-zeta) (1 + lambda cos_ijk)^zeta
//                 exp(-eta (r_ij^2 + r_ik^2 + r_jk^2)) fc(r_ij) fc(r_ik) fc(r_jk)
//   fc(r)     = (cos(pi r / Rc) + 1) / 2, pairs and the opposite edge < Rc,
//               minimum image from Domain::shortest
//   E_i       = W23 tanh(W12 tanh(W01 (G o range) + b01) + b12) + b23, in eV,
//               scaled by 10 eV Na to kJ/mol
//   f_m       = -sum_i dE_i/dG_i . dG_i/dr_m
//
// Device formulation: the reference gathers, for every atom, the derivatives
// of all neighbours' descriptors (about 1.5 n^2 triplets per atom).  Here
// each centre i is one wavefront that stages its in-range neighbours in LDS,
// and
//   pass 1 (k_bp_desc)   sums G_i over pairs j<k (n^2/2 triplets);
//   pass 2 (k_bp_mlp)    runs the per-element MLP forward and backward for
//                        dE_i/dG_i on MFMA (FP16 inputs split into hi + lo
//                        halves, three products per tile, FP32 accumulation),
//                        or in FP64 on the vector units in FP64 mode;
//   pass 3 (k_bp_grad)   has lane j sum d(w_i . G_i)/dr_j over all k (ordered
//                        pairs, no atomics), giving the force that centre i
//                        puts on each neighbour slot, and the centre's own
//                        share by translation invariance;
//   pass 4 (k_bp_gather) adds, for every atom, its own share and the shares
//                        its neighbours put on it (through a reverse-slot
//                        index built at list rebuilds), then runs the fused
//                        velocity-Verlet update.
#include <cmath>
#include <cstring>
#include <fstream>
#include <functional>
#include <memory>
#include <sstream>
#include <string>
#include <vector>
#include "kern.h"

namespace mdc {

constexpr int kBpMaxR = 16, kBpMaxA = 64, kBpMaxEta = 8, kBpMaxHidden = 64, kBpMaxIn = 64;
constexpr int kBpMaxG = 8;  // distinct (lambda, zeta) pairs of the G4 terms
// Ladder layout: every G4 term has lambda = +-1 and zeta = 2^b, b < 5.  The
// groups are then g = 2 b + (lambda < 0), all 2 lz of them (lz = 1 + the
// largest b), and (1 + lambda cos)^zeta comes from repeated squaring of
// 1 + cos and 1 - cos without a per-group power.
constexpr int kBpMaxLadder = 5;
constexpr int kBpGridG = 2 * kBpMaxLadder;  // group rows of the term grid

// Rungs of the ladder kernels that serve nr radial terms, ne G4 etas and
// zeta up to 2^(lz - 1) (rungs without terms are summed and dropped), or 0
// when no ladder kernel does: 8 G2 with 2 etas to zeta 4 (the 16-input Cu
// model) and 8 G2 with 7 etas to zeta 16 (the 51-input Cu model).
inline int ladder_rungs(int nr, int ne, int lz) {
    if (nr == 8 && ne == 2 && lz <= 3) return 3;
    if (nr == 8 && ne == 7 && lz <= 5) return 5;
    return 0;
}

struct BpDesc {
    int nr, na, nin, h1, h2, ne, ng;
    double Rc, Rc2, Rs, pi_rc, escale;
    double eta_r[kBpMaxR];
    double eta_e[kBpMaxEta];
    int a_e[kBpMaxA];
    double a_coef[kBpMaxA], a_lam[kBpMaxA], a_zeta[kBpMaxA];
    int a_iz[kBpMaxA];
    // G4 terms on a (group, eta) grid: group g has lambda g_lam, zeta g_zeta
    // (g_iz = integer zeta or -1); grid term t = g * kBpMaxEta + e is angular
    // input t_q[t] (-1: none) with coefficient t_coef[t]
    double g_lam[kBpGridG], g_zeta[kBpGridG], g_zl[kBpGridG];
    int g_iz[kBpGridG];
    int t_q[kBpGridG * kBpMaxEta];
    double t_coef[kBpGridG * kBpMaxEta];
    int lz;  // ladder layout: the number of zeta rungs; 0: groups in order of appearance
    // species (1 or 2) and the block offsets of the input vector
    // (tools/bpnn/maise_model.cxx:113-131): radial block sr[d] of the kind
    // difference d = [k_j != k_i], angular block sa[d_ij + d_ik]
    int nspc;
    int sr[2], sa[3];
};

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

struct BpnnParams {
    BpDesc h{};
    BpDesc* d = nullptr;
    double* w64 = nullptr;  // W01 [h1][nin] (range-scaled), b01, W12 [h2][h1], b12, W23 [h2], b23
    float* w32 = nullptr;   // b01, b12, W23, b23 as float for the MFMA epilogues
    float* w64f = nullptr;  // w64 in FP32 (same layout and stride): the fused pass's network
    half4_t* bsplit = nullptr;
    int boff[4] = {0, 0, 0, 0};
    int nt[4] = {0, 0, 0, 0}, kt[4] = {0, 0, 0, 0};
    int nin_pad = 0, h1_pad = 0, h2_pad = 0;
    // elements of one species' block in w64, w32 and bsplit
    size_t w64_stride = 0, w32_stride = 0, b_stride = 0;
};

// Work arrays of one BPNN evaluation: descriptors G and dE/dG w [n][ldg],
// force shares P [n][cap][3] and the centres' own shares own [n][3], the
// virial shares PV [n][cap][3] and OV [n][3] (only when the virial is
// asked for), and the reverse-slot index rev (one int per list element).
struct BpBufs {
    void* G = nullptr;
    void* w = nullptr;
    void* P = nullptr;
    void* own = nullptr;
    void* PV = nullptr;
    void* OV = nullptr;
    int* rev = nullptr;
};

// per-batch BPNN work arrays (in Batch::bp_scratch), sized for the list
// they were made for: rebuilt when the capacity, the atom count, the
// precision or the list's element count (which grows when the lanes per
// atom drop) no longer match
struct BpWork {
    int cap = 0, n = 0, prec = -1;
    size_t rev_elems = 0;
    bool has_vir = false;
    BpBufs a;
    bool rev_valid = false;
    int rev_builds = -1;
};

// --------------------------------------------------------------------------
// device helpers
// --------------------------------------------------------------------------
__device__ inline float dcos_(float x) { return __cosf(x); }
__device__ inline double dcos_(double x) { return cos(x); }
__device__ inline float dsin_(float x) { return __sinf(x); }
__device__ inline double dsin_(double x) { return sin(x); }
__device__ inline float dexp_(float x) { return __expf(x); }
__device__ inline double dexp_(double x) { return exp(x); }
__device__ inline float dsqrt_(float x) { return sqrtf(x); }
__device__ inline double dsqrt_(double x) { return sqrt(x); }
// 1/x and 1/sqrt(x): the hardware approximations (1 ulp) in FP32, exact in FP64
__device__ inline float drcp_(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline double drcp_(double x) { return 1.0 / x; }
__device__ inline float drsq_(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ inline double drsq_(double x) { return 1.0 / sqrt(x); }

// b^k for a whole k >= 0 (uniform across the wavefront), by squaring
template <typename R>
__device__ inline R powu_(R b, int k) {
    R p = (R)1;
    for (; k > 0; k >>= 1) {
        if (k & 1) p *= b;
        b *= b;
    }
    return p;
}

template <typename R>
__device__ inline R powi_(R b, int iz, R z) {
    if (iz >= 0) {
        R p = (R)1;
        for (int q = 0; q < iz; ++q) p *= b;
        return p;
    }
    return (R)pow((double)b, (double)z);
}

template <typename R>
__device__ inline R wave_sum(R v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// The FP64 descriptor pass compacts the neighbour pairs in the cutoff
// before their arithmetic (about half of all pairs of a centre's
// neighbours are farther apart than the cutoff): it queues them per wave,
// kPairQueue entries (up to 63 waiting and 64 new).  In FP32 the queue's
// own instructions cost about what it saves.
constexpr int kPairQueue = 2 * kWave;

template <typename R>
struct Stage {
    R *x, *y, *z, *d, *id, *fc, *dfc, *E;  // E: [ne][cap]; id = 1/d
    R* g;                                  // force-share accumulators [3][cap] (pass 3)
    int* slot;
    int* queue;  // descriptor pass: in-range pairs a | b << 16
    int cap;
};

// grad: the force pass also keeps the force-share accumulators and each
// staged neighbour's list slot; the descriptor pass stages without them
// (less LDS per wavefront, more wavefronts per CU).
template <typename R>
__device__ inline Stage<R> carve(char* base, int cap, int ne, bool grad) {
    Stage<R> S;
    R* p = reinterpret_cast<R*>(base);
    int k = 0;
    auto take = [&](int rows) {
        R* q = p + (size_t)k * cap;
        k += rows;
        return q;
    };
    S.x = take(1);
    S.y = take(1);
    S.z = take(1);
    S.d = take(1);
    S.fc = take(1);
    S.dfc = take(1);
    S.id = take(1);
    S.E = take(ne);
    S.g = grad ? take(3) : nullptr;
    int* tail = reinterpret_cast<int*>(p + (size_t)k * cap);
    S.slot = grad ? tail : nullptr;
    S.queue = grad ? nullptr : tail;
    S.cap = cap;
    return S;
}

__host__ __device__ inline size_t stage_bytes(int cap, int ne, int rsize, bool grad) {
    const int rows = 7 + ne + (grad ? 3 : 0);
    const size_t queue = grad || rsize != 8 ? 0 : (size_t)kPairQueue * sizeof(int);
    // a multiple of 8 bytes, so that every wavefront's stage starts aligned
    // for FP64 rows whatever the capacity
    return ((size_t)cap * (rows * rsize + (grad ? 4 : 0)) + queue + 7) & ~(size_t)7;
}

// First unordered pair (a, b = a + 1 + t) of the flattened pair index t
// over nc staged neighbours; advance() moves to index t + kWave.
struct PairWalk {
    int a, t, nc;
    __device__ PairWalk(int nc_, int t0) : a(0), t(t0), nc(nc_) { settle(); }
    __device__ void settle() {
        while (a < nc - 1 && t >= nc - 1 - a) {
            t -= nc - 1 - a;
            ++a;
        }
    }
    __device__ bool valid() const { return a < nc - 1; }
    __device__ int b() const { return a + 1 + t; }
    __device__ void advance() {
        t += kWave;
        settle();
    }
};

// Every unordered pair of n staged neighbours once, as (a, b = a + d mod
// n): d = 1 .. (n - 1) / 2 for every a, and d = n / 2 for a < n / 2 when n
// is even.  The pairs are taken in order of d, then a (index t = (d - 1) n
// + a), 64 at a time, lane l at step k on index 64 k + l: the lanes of a
// step hold consecutive a and consecutive b, so their LDS reads of both
// neighbours' values fall in different banks, and a lane moves on by adding
// 64 to its index (a += 64 mod n, d += 64 div n and the carry), with no
// search.  fn(a, b, live) runs at every step in every lane; live is false
// past the last pair, where (a, b) is a pair of staged neighbours to be
// ignored.
template <class Fn>
__device__ inline void pair_walk(int n, int lane, Fn&& fn) {
    if (n < 2) return;
    const int hd = (n - 1) / 2;
    const int ex = (n % 2 == 0) ? n / 2 : 0;  // rows with one more partner (d = n / 2)
    const int T = n * hd + ex;
    const int steps = (T + kWave - 1) / kWave;
    const int q = kWave / n, r = kWave - q * n;  // 64 = q n + r
    int a = lane % n, d = lane / n + 1;
    for (int k = 0, t = lane; k < steps; ++k, t += kWave) {
        int b = a + d;
        if (b >= n) b -= n;
        fn(a, min(b, n - 1), t < T);
        a += r;
        d += q;
        if (a >= n) {
            a -= n;
            ++d;
        }
    }
}

// Stage the in-range neighbours of centre i (1e-15 <= r^2 < Rc^2 under
// Domain::shortest) into LDS, compacted in list order; out-of-range slots
// get zero force shares in Pzero and zero virial shares in PVzero when
// those are given.  ck >= 0 (two species, ck the centre's kind): the
// neighbours of kind ck go to the front [0, returned count), the others to
// the back [cap - nd, cap).
template <typename R>
__device__ int bp_stage(int i, const double4& pi, const SysDesc& s, const int* __restrict__ nl, int cnt, int tpa,
                        int M, const double4* __restrict__ pos, const BpDesc& P, const Stage<R>& S,
                        R* __restrict__ Pzero, R* __restrict__ PVzero, int cap, int ck, int& nd) {
    const int lane = threadIdx.x & 63;
    int base = 0, based = 0;
    for (int k0 = 0; k0 < cnt; k0 += kWave) {
        const int k = k0 + lane;
        bool ok = false, other = false;
        double dx = 0, dy = 0, dz = 0, d2 = 0;
        if (k < cnt) {
            const int j = ent_j(nl[list_index(i, k, tpa, M)]);
            const double4 pj = pos[j];
            other = ck >= 0 && w_kind(pj.w) != ck;
            dx = pj.x - pi.x;
            dy = pj.y - pi.y;
            dz = pj.z - pi.z;
            shortest3(s, dx, dy, dz);
            d2 = sq3_exact(dx, dy, dz);
            ok = d2 >= 1e-15 && d2 < P.Rc2;
            if (!ok && Pzero) {
                R* q = Pzero + ((size_t)i * cap + k) * 3;
                q[0] = q[1] = q[2] = (R)0;
            }
            if (!ok && PVzero) {
                R* q = PVzero + ((size_t)i * cap + k) * 3;
                q[0] = q[1] = q[2] = (R)0;
            }
        }
        const unsigned long long m = __ballot(ok && !other), md = __ballot(ok && other);
        if (ok) {
            const unsigned long long lt = (1ull << lane) - 1ull;
            const int at = other ? cap - 1 - (based + __popcll(md & lt)) : base + __popcll(m & lt);
            const double d = sqrt(d2);
            S.x[at] = (R)dx;
            S.y[at] = (R)dy;
            S.z[at] = (R)dz;
            S.d[at] = (R)d;
            S.id[at] = (R)(1.0 / d);
            // cutoff function and its derivative, exp(-eta d^2), in the
            // precision of the pass
            const R rd = (R)d, prc = (R)P.pi_rc;
            S.fc[at] = (R)0.5 * (dcos_(prc * rd) + (R)1);
            S.dfc[at] = (R)-0.5 * dsin_(prc * rd) * prc;
            for (int e = 0; e < P.ne; ++e) S.E[e * S.cap + at] = dexp_(-(R)P.eta_e[e] * (R)d2);
            if (S.slot) S.slot[at] = k;
        }
        base += __popcll(m);
        based += __popcll(md);
    }
    // LDS traffic of one wavefront is processed in order; the barrier only
    // stops the compiler from moving the readers above the staging stores
    __builtin_amdgcn_wave_barrier();
    nd = based;
    return base;
}

// pass 1: descriptors.  Per pair of staged neighbours the G4 factor
// (1 + lambda cos)^zeta is formed once per (lambda, zeta) group and shared by
// the group's eta terms; NR, NG, NE bound the radial, group and eta counts.
// LZ > 0: the ladder layout with LZ rungs (NG = 2 LZ), NR and NE exact.
// K2: two species; the neighbours of the centre's kind are staged first
// and those of the other kind last, the radial terms go to the block of the
// kind difference and the angular terms of the same/same, same/other and
// other/other pairs to blocks 0, 1 and 2 (maise_bpnn.cxx:141-257 for two
// kinds, whose differences are 0 or 1).
// The descriptors of one centre from its staged neighbours, by one
// wavefront: the radial and angular sums in acc (every lane holds the
// totals on return, single species), written to grow[0, ldg) when grow is
// given (two species: always, block by block).
template <typename R, int NR, int NG, int NE, int LZ, bool K2, int ZM = -1>
__device__ inline void desc_wave(const BpDesc& P, const Stage<R>& S, int nc, int nd, int cap, int lane, R* grow,
                                 int ldg, R (&acc)[NR + NG * NE]) {
    const int d0 = cap - nd;  // the other kind's neighbours: [d0, cap)
    R acc1[K2 ? NR : 1];  // radial block of the other kind
#pragma unroll
    for (int q = 0; q < NR + NG * NE; ++q) acc[q] = (R)0;
#pragma unroll
    for (int q = 0; q < (K2 ? NR : 1); ++q) acc1[q] = (R)0;
    if constexpr (K2) {
        // blocks are stored as they complete
        for (int q = lane; q < ldg; q += kWave) grow[q] = (R)0;
        __builtin_amdgcn_wave_barrier();
    }
    // descriptor parameters in registers, in the precision of the pass
    const int nr = P.nr, ng = P.ng, ne = P.ne;
    R eta_r[NR], g_lam[NG], eta_e[NE];
    int g_iz[NG];
#pragma unroll
    for (int r = 0; r < NR; ++r) eta_r[r] = r < nr ? (R)P.eta_r[r] : (R)0;
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        g_lam[g] = g < ng ? (R)P.g_lam[g] : (R)0;
        g_iz[g] = g < ng ? P.g_iz[g] : 0;
    }
#pragma unroll
    for (int e = 0; e < NE; ++e) eta_e[e] = e < ne ? (R)P.eta_e[e] : (R)0;
    const R Rs = (R)P.Rs;
    for (int a = lane; a < nc; a += kWave) {
        const R d = S.d[a], fc = S.fc[a];
#pragma unroll
        for (int r = 0; r < NR; ++r)
            if (LZ > 0 || r < nr) acc[r] += dexp_(-eta_r[r] * (d - Rs) * (d - Rs)) * fc;
    }
    if constexpr (K2)
        for (int a = d0 + lane; a < cap; a += kWave) {
            const R d = S.d[a], fc = S.fc[a];
#pragma unroll
            for (int r = 0; r < NR; ++r)
                if (LZ > 0 || r < nr) acc1[r] += dexp_(-eta_r[r] * (d - Rs) * (d - Rs)) * fc;
        }
    // pairs a < b over the staged list: the flattened pair index strided
    // over the lanes is tested against the cutoff; in FP64 the pairs in
    // range are queued (ballot and prefix count) and taken 64 at a time, so
    // that every lane of the arithmetic has a pair
    const R Rc2 = (R)P.Rc2, pirc = (R)P.pi_rc;
    // a pair's angular terms from its cosine and its exp / cutoff factors
    auto ang_terms = [&](R cs, const R* Ex) {
        if constexpr (LZ > 0) {
            // rung b: (1 + cos)^(2^b) and (1 - cos)^(2^b) by squaring
            R pp = (R)1 + cs, pm = (R)1 - cs;
#pragma unroll
            for (int b = 0; b < LZ; ++b) {
                if ((ZM >> b) & 1)
#pragma unroll
                    for (int e = 0; e < NE; ++e) {
                        acc[NR + (2 * b) * NE + e] += pp * Ex[e];
                        acc[NR + (2 * b + 1) * NE + e] += pm * Ex[e];
                    }
                if (b + 1 < LZ) {
                    pp *= pp;
                    pm *= pm;
                }
            }
        } else {
#pragma unroll
            for (int g = 0; g < NG; ++g)
                if (g < ng) {
                    const R pz = powu_((R)1 + g_lam[g] * cs, g_iz[g]);
#pragma unroll
                    for (int e = 0; e < NE; ++e) acc[NR + g * NE + e] += pz * Ex[e];
                }
        }
    };
    auto pair_terms = [&](int a, int b) {
        const R cx = S.x[b] - S.x[a], cy = S.y[b] - S.y[a], cz = S.z[b] - S.z[a];
        const R d2c = cx * cx + cy * cy + cz * cz;
        const R dc = d2c * drsq_(d2c);
        const R fcc = (R)0.5 * (dcos_(pirc * dc) + (R)1);
        const R cs = (S.x[a] * S.x[b] + S.y[a] * S.y[b] + S.z[a] * S.z[b]) * S.id[a] * S.id[b];
        const R F = S.fc[a] * S.fc[b] * fcc;
        R Ex[NE];
#pragma unroll
        for (int e = 0; e < NE; ++e)
            Ex[e] = (LZ > 0 || e < ne) ? S.E[e * cap + a] * S.E[e * cap + b] * dexp_(-eta_e[e] * d2c) * F : (R)0;
        ang_terms(cs, Ex);
    };
    // FP32 walks: the values of the pair's row a and of b, read at each step
    R rx = 0, ry = 0, rz = 0, rid = 0, rfc = 0, rE[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) rE[e] = (R)0;
    auto load_row = [&](int a) {
        rx = S.x[a];
        ry = S.y[a];
        rz = S.z[a];
        rid = S.id[a];
        rfc = S.fc[a];
#pragma unroll
        for (int e = 0; e < NE; ++e) rE[e] = (LZ > 0 || e < ne) ? S.E[e * cap + a] : (R)0;
    };
    // all of b's values are read at once, and a pair out of range (or a
    // lane past its run) gets F = 0 by a 0/1 factor rather than a branch:
    // the lanes of a step run the arithmetic whenever any of them has a
    // pair in range, and a branch only adds the copies that merge its two
    // sides.  1/|r_c| is taken of at least 1e-30 so that every factor is
    // finite.
    auto pair_terms_row = [&](int b, bool live) {
        const R bx = S.x[b], by = S.y[b], bz = S.z[b], bid = S.id[b], bfc = S.fc[b];
        R bE[NE];
#pragma unroll
        for (int e = 0; e < NE; ++e) bE[e] = (LZ > 0 || e < ne) ? S.E[e * cap + b] : (R)0;
        const R cx = bx - rx, cy = by - ry, cz = bz - rz;
        const R d2c = cx * cx + cy * cy + cz * cz;
        const R ok = (live && d2c >= (R)1e-15 && d2c < Rc2) ? (R)1 : (R)0;
        const R dc = d2c * drsq_(fmax(d2c, (R)1e-30));
        const R fcc = (R)0.5 * (dcos_(pirc * dc) + (R)1);
        const R cs = (rx * bx + ry * by + rz * bz) * rid * bid;
        const R F = ok * rfc * bfc * fcc;
        R Ex[NE];
#pragma unroll
        for (int e = 0; e < NE; ++e)
            Ex[e] = (LZ > 0 || e < ne) ? rE[e] * bE[e] * dexp_(-eta_e[e] * d2c) * F : (R)0;
        ang_terms(cs, Ex);
    };
    if constexpr (sizeof(R) == 8) {
        int* Q = S.queue;
        int qn = 0;  // queued pairs (uniform over the wave)
        PairWalk pw(nc, lane);
        while (true) {
            bool ok = false;
            int ent = 0;
            if (pw.valid()) {
                const int a = pw.a, b = pw.b();
                const R cx = S.x[b] - S.x[a], cy = S.y[b] - S.y[a], cz = S.z[b] - S.z[a];
                const R d2c = cx * cx + cy * cy + cz * cz;
                ok = d2c >= (R)1e-15 && d2c < Rc2;
                ent = a | (b << 16);
                pw.advance();
            }
            const unsigned long long m = __ballot(ok);
            if (ok) Q[qn + __popcll(m & ((1ull << lane) - 1ull))] = ent;
            qn += __popcll(m);
            const bool more = __ballot(pw.valid()) != 0ull;
            __builtin_amdgcn_wave_barrier();
            while (qn >= kWave || (!more && qn > 0)) {
                const int take = qn < kWave ? qn : kWave;
                qn -= take;
                if (lane < take) {
                    const int e = Q[qn + lane];
                    pair_terms(e & 0xffff, e >> 16);
                }
                __builtin_amdgcn_wave_barrier();
            }
            if (!more) break;
        }
    } else {
        pair_walk(nc, lane, [&](int a, int b, bool live) {
            load_row(a);
            pair_terms_row(b, live);
        });
    }
    if constexpr (K2) {
        // the angular sums of the pairs walked so far into block `key`
        auto store_ang = [&](int key) {
#pragma unroll
            for (int q = NR; q < NR + NG * NE; ++q) acc[q] = wave_sum(acc[q]);
            if (lane == 0) {
#pragma unroll
                for (int gg = 0; gg < NG; ++gg)
#pragma unroll
                    for (int e = 0; e < NE; ++e) {
                        const int t = gg * kBpMaxEta + e;
                        if (gg < P.ng && e < P.ne && P.t_q[t] >= 0)
                            grow[P.sa[key] + P.t_q[t]] = (R)P.t_coef[t] * acc[NR + gg * NE + e];
                    }
            }
#pragma unroll
            for (int q = NR; q < NR + NG * NE; ++q) acc[q] = (R)0;
        };
        auto in_range = [&](int a, int b) {
            const R cx = S.x[b] - S.x[a], cy = S.y[b] - S.y[a], cz = S.z[b] - S.z[a];
            const R d2c = cx * cx + cy * cy + cz * cz;
            return d2c >= (R)1e-15 && d2c < Rc2;
        };
        store_ang(0);
        // same/other pairs: a in [0, nc), b in [d0, cap)
        for (int t = lane; t < nc * nd; t += kWave) {
            const int a = t / nd, b = d0 + (t - a * nd);
            if (in_range(a, b)) pair_terms(a, b);
        }
        store_ang(1);
        if constexpr (sizeof(R) == 8) {
            for (PairWalk pw(nd, lane); pw.valid(); pw.advance()) {
                const int a = d0 + pw.a, b = d0 + pw.b();
                if (in_range(a, b)) pair_terms(a, b);
            }
        } else {
            pair_walk(nd, lane, [&](int a, int b, bool live) {
                load_row(d0 + a);
                pair_terms_row(d0 + b, live);
            });
        }
        store_ang(2);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            acc[r] = wave_sum(acc[r]);
            acc1[r] = wave_sum(acc1[r]);
        }
        if (lane == 0)
#pragma unroll
            for (int r = 0; r < NR; ++r)
                if (r < P.nr) {
                    grow[P.sr[0] + r] = acc[r];
                    grow[P.sr[1] + r] = acc1[r];
                }
        return;
    }
#pragma unroll
    for (int q = 0; q < NR + NG * NE; ++q) acc[q] = wave_sum(acc[q]);
    if (grow && lane == 0) {
        R* g = grow;
        for (int q = 0; q < ldg; ++q) g[q] = (R)0;
#pragma unroll
        for (int r = 0; r < NR; ++r)
            if (r < P.nr) g[r] = acc[r];
#pragma unroll
        for (int gg = 0; gg < NG; ++gg)
#pragma unroll
            for (int e = 0; e < NE; ++e) {
                const int t = gg * kBpMaxEta + e;
                if (gg < P.ng && e < P.ne && P.t_q[t] >= 0) g[P.nr + P.t_q[t]] = (R)P.t_coef[t] * acc[NR + gg * NE + e];
            }
    }
}

template <typename R, int NR, int NG, int NE, int LZ = 0, bool K2 = false>
__global__ void __launch_bounds__(256) k_bp_desc(const double4* __restrict__ pos, const int* __restrict__ nl,
                                                 const int* __restrict__ nn, int tpa, int M, int cap,
                                                 const SysDesc* __restrict__ desc, const BpDesc* __restrict__ Pd,
                                                 R* __restrict__ G, int ldg, int n, StepArgs SA,
                                                 int* __restrict__ misc) {
    if (skip_step(SA, misc)) return;
    extern __shared__ char smem[];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + wv;
    if (i >= n) return;
    const BpDesc& P = *Pd;
    const Stage<R> S = carve<R>(smem + wv * stage_bytes(cap, P.ne, sizeof(R), false), cap, P.ne, false);
    const double4 pi = pos[i];
    const SysDesc& s = desc[w_sys(pi.w)];
    const int cnt = min(nn[i], cap);
    int nd = 0;
    const int nc = bp_stage<R>(i, pi, s, nl, cnt, tpa, M, pos, P, S, (R*)nullptr, (R*)nullptr, cap,
                               K2 ? (int)w_kind(pi.w) : -1, nd);
    R acc[NR + NG * NE];
    desc_wave<R, NR, NG, NE, LZ, K2>(P, S, nc, nd, cap, lane, G + (size_t)i * ldg, ldg, acc);
}

// pass 2 (FP64): L lanes per atom (32 when the input and hidden widths fit,
// else 64), lane q holding unit q of each layer; activations move between
// lanes by shuffles, the weights are read row- or column-wise so a group's
// reads are contiguous (W01T and W12T are transposed copies).
template <int L>
__global__ void __launch_bounds__(256) k_bp_mlp64(const double* __restrict__ G, int ldg, const double* __restrict__ W,
                                                  const BpDesc* __restrict__ Pd, const double4* __restrict__ pos,
                                                  double* __restrict__ wout, double* __restrict__ Ep, int n, int sp,
                                                  StepArgs SA, int* __restrict__ misc) {
    // sp >= 0: the network of species sp, for the atoms of that kind only
    if (skip_step(SA, misc)) return;
    const int lane = threadIdx.x & 63;
    const int sub = lane & (L - 1), base = lane - sub;
    const int i = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) / L);
    const bool valid = i < n;
    const BpDesc& P = *Pd;
    const int nin = P.nin, h1 = P.h1, h2 = P.h2;
    const double* W01 = W;
    const double* b01 = W01 + h1 * nin;
    const double* W12 = b01 + h1;
    const double* b12 = W12 + h2 * h1;
    const double* W23 = b12 + h2;
    const double b23 = W23[h2];
    const double* W01T = W23 + h2 + 1;  // [nin][h1]
    const double* W12T = W01T + nin * h1;  // [h1][h2]
    const bool mine = valid && (sp < 0 || w_kind(pos[i].w) == sp);
    const bool act = mine && !w_inactive(pos[i].w);
    const double* g = G + (size_t)(valid ? i : 0) * ldg;
    // every lane of the wave takes part in every shuffle below
    double y1 = 0.0;
    if (act && sub < h1) {
        double s = 0.0;
        for (int k = 0; k < nin; ++k) s += W01T[k * h1 + sub] * g[k];
        y1 = tanh(s + b01[sub]);
    }
    double s2 = 0.0;
    for (int p = 0; p < h1; ++p) {
        const double yp = __shfl(y1, base + p, kWave);
        if (sub < h2) s2 += W12T[p * h2 + sub] * yp;
    }
    double t = 0.0, e = 0.0;
    if (act && sub < h2) {
        const double y2 = tanh(s2 + b12[sub]);
        e = W23[sub] * y2;
        t = W23[sub] * (1.0 - y2 * y2);
    }
    for (int o = L >> 1; o > 0; o >>= 1) e += __shfl_xor(e, o, kWave);
    double s1 = 0.0;
    for (int q = 0; q < h2; ++q) {
        const double tq = __shfl(t, base + q, kWave);
        if (sub < h1) s1 += W12[q * h1 + sub] * tq;
    }
    const double d1 = sub < h1 ? s1 * (1.0 - y1 * y1) : 0.0;
    double so = 0.0;
    for (int q = 0; q < h1; ++q) {
        const double dq = __shfl(d1, base + q, kWave);
        if (sub < nin) so += W01[q * nin + sub] * dq;
    }
    if (!mine) return;
    double* wo = wout + (size_t)i * ldg;
    for (int k = sub; k < ldg; k += L) wo[k] = (act && k < nin) ? so * P.escale : 0.0;
    if (act && sub == 0) Ep[i] = (b23 + e) * P.escale;
}

// pass 2 (mixed): one wavefront per 16 atoms on v_mfma_f32_16x16x16_f16.
// Operand layout of the 16x16x16 instruction: lane l holds A[l%16][4(l/16)+c]
// and B[4(l/16)+c][l%16]; the result D[4(l/16)+c][l%16].  Each FP32 operand
// x is split into hi = fp16(x), lo = fp16(x - hi), and a tile product is
// A_hi B_hi + A_hi B_lo + A_lo B_hi.
__device__ inline void split_h(float x, _Float16& hi, _Float16& lo) {
    hi = (_Float16)x;
    lo = (_Float16)(x - (float)hi);
}

// C[16 x nt*16] = A[16 x kt*16] (LDS, row stride lda) * B (pre-split lane
// layout) ; result tile nt written to LDS out[row][nt*16 + col]
__device__ inline void mfma_gemm(const float* As, int lda, const half4_t* __restrict__ B, int nt, int kt, float* out,
                                 int ldo, const float* bias, int act_tanh, int ncols) {
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, kq = (lane >> 4) * 4;
    for (int t = 0; t < nt; ++t) {
        float4_t acc = {0.f, 0.f, 0.f, 0.f};
        for (int k = 0; k < kt; ++k) {
            half4_t ahi, alo;
            for (int c = 0; c < 4; ++c) {
                _Float16 h, l;
                split_h(As[r * lda + k * 16 + kq + c], h, l);
                ahi[c] = h;
                alo[c] = l;
            }
            const half4_t bhi = B[((size_t)(t * kt + k) * 2 + 0) * 64 + lane];
            const half4_t blo = B[((size_t)(t * kt + k) * 2 + 1) * 64 + lane];
            acc = __builtin_amdgcn_mfma_f32_16x16x16f16(ahi, bhi, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x16f16(ahi, blo, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x16f16(alo, bhi, acc, 0, 0, 0);
        }
        const int col = t * 16 + (lane & 15);
        for (int c = 0; c < 4; ++c) {
            const int row = (lane >> 4) * 4 + c;
            float v = acc[c];
            if (bias) v += col < ncols ? bias[col] : 0.f;
            if (act_tanh) v = col < ncols ? tanhf(v) : 0.f;
            out[row * ldo + col] = v;
        }
    }
}

__global__ void __launch_bounds__(64) k_bp_mlp_mfma(const float* __restrict__ G, int ldg, const half4_t* __restrict__ B,
                                                    int4 boff, int4 nts, int4 kts, const float* __restrict__ w32,
                                                    const BpDesc* __restrict__ Pd, const double4* __restrict__ pos,
                                                    float* __restrict__ wout, double* __restrict__ Ep, int n, int sp,
                                                    StepArgs SA, int* __restrict__ misc) {
    // sp >= 0: the network of species sp, for the atoms of that kind only
    if (skip_step(SA, misc)) return;
    __shared__ float X[16 * kBpMaxHidden], H1[16 * kBpMaxHidden], H2[16 * kBpMaxHidden], T[16 * kBpMaxHidden];
    auto mine = [&](int i) { return i < n && (sp < 0 || w_kind(pos[i].w) == sp); };
    const BpDesc& P = *Pd;
    const int lane = threadIdx.x;
    const int a0 = blockIdx.x * 16;
    const int nin = P.nin, h1 = P.h1, h2 = P.h2;
    const int inp = kts.x * 16, h1p = nts.x * 16, h2p = nts.y * 16;
    const float* b01 = w32;
    const float* b12 = b01 + h1;
    const float* W23 = b12 + h2;
    const float b23 = W23[h2];
    const float* xs = W23 + h2 + 1;  // power-of-two input scales [nin_pad]
    for (int e = lane; e < 16 * inp; e += 64) {
        const int r = e / inp, c = e % inp;
        const int i = a0 + r;
        X[r * kBpMaxHidden + c] = (i < n && c < nin) ? G[(size_t)i * ldg + c] * xs[c] : 0.f;
    }
    __syncthreads();
    // forward
    mfma_gemm(X, kBpMaxHidden, B + boff.x, nts.x, kts.x, H1, kBpMaxHidden, b01, 1, h1);
    __syncthreads();
    mfma_gemm(H1, kBpMaxHidden, B + boff.y, nts.y, kts.y, H2, kBpMaxHidden, b12, 1, h2);
    __syncthreads();
    if (lane < 16) {
        const int i = a0 + lane;
        float E = b23;
        for (int q = 0; q < h2; ++q) E += W23[q] * H2[lane * kBpMaxHidden + q];
        if (mine(i) && !w_inactive(pos[i].w)) Ep[i] = (double)E * P.escale;
    }
    // backward: D2 = W23 o (1 - h2^2)
    for (int e = lane; e < 16 * h2p; e += 64) {
        const int r = e / h2p, c = e % h2p;
        const float y = H2[r * kBpMaxHidden + c];
        T[r * kBpMaxHidden + c] = c < h2 ? W23[c] * (1.f - y * y) : 0.f;
    }
    __syncthreads();
    // D1 = (D2 W12) o (1 - h1^2): B[k][n] = W12[k][n]
    mfma_gemm(T, kBpMaxHidden, B + boff.z, nts.z, kts.z, H2, kBpMaxHidden, nullptr, 0, h1);
    __syncthreads();
    for (int e = lane; e < 16 * h1p; e += 64) {
        const int r = e / h1p, c = e % h1p;
        const float y = H1[r * kBpMaxHidden + c];
        T[r * kBpMaxHidden + c] = c < h1 ? H2[r * kBpMaxHidden + c] * (1.f - y * y) : 0.f;
    }
    __syncthreads();
    // dE/dG = D1 W01: B[k][n] = W01[k][n]
    mfma_gemm(T, kBpMaxHidden, B + boff.w, nts.w, kts.w, X, kBpMaxHidden, nullptr, 0, nin);
    __syncthreads();
    for (int e = lane; e < 16 * ldg; e += 64) {
        const int r = e / ldg, c = e % ldg;
        const int i = a0 + r;
        if (mine(i)) {
            const bool act = !w_inactive(pos[i].w);
            wout[(size_t)i * ldg + c] =
                (act && c < nin) ? (float)(X[r * kBpMaxHidden + c] * xs[c] * P.escale) : 0.f;
        }
    }
}

// pass 3: force shares of centre i on its neighbour slots and on itself.
// VIR: also the diagonal virial of maise_bpnn.cxx:98-120.  Its radial term
// on atom m sums, over the centres i listing m, d(w_i . G2_i)/dr r_im(c)^2
// / r_im with the centre's dE/dG: centre i writes that share for each slot
// (PV, gathered like the force shares).  Its angular term on atom i is the
// centre's own dE/dG contracted with the pattern_dG4_s_ stress of
// tools/bpnn/base.h:157-181 over the pairs of i's neighbours (OV).
// The pair's sums over the G4 groups of the gradient pass: with c_q =
// dE/dG_q coef_q and Ex_e = E_e(a) E_e(b) exp(-eta_e d_c^2),
// SA = sum_q c_q Ex pz_q, SAe = sum_q c_q eta Ex pz_q and
// SZ = sum_q c_q Ex zeta lam pz1_q, where pz = (1 + lam cos)^zeta and pz1 =
// (1 + lam cos)^(zeta - 1).  LZ > 0 (ladder layout): rung b holds
// (1 +- cos)^(2^b) and (1 +- cos)^(2^b - 1) by squaring and running
// products; the rungs are summed per eta first, A_e = sum_b (c+ P+ + c- P-)
// and Z_e = sum_b (cz+ Q+ + cz- Q-) with Cz the coefficients times 2^b and
// -2^b, then SA = sum_e Ex_e A_e, SAe = sum_e eta_e Ex_e A_e, SZ = sum_e
// Ex_e Z_e.  Rungs outside the mask ZM have no terms and are skipped (their
// powers are still formed for the rungs above).
template <typename R, int NG, int NE, int LZ, int ZM = -1>
__device__ inline void group_sums(R cs, const R* C, const R* Cz, const R* Ex, const R* Exe, const R* g_lam,
                                  const R* g_zl, const int* g_iz, int ng, R& SA, R& SAe, R& SZ) {
    if constexpr (LZ > 0) {
        R Pp = (R)1 + cs, Pm = (R)1 - cs, Qp = (R)1, Qm = (R)1;
        R A[NE], Z[NE];
#pragma unroll
        for (int e = 0; e < NE; ++e) A[e] = Z[e] = (R)0;
#pragma unroll
        for (int b = 0; b < LZ; ++b) {
            if ((ZM >> b) & 1) {
#pragma unroll
                for (int e = 0; e < NE; ++e) {
                    A[e] += C[(2 * b) * NE + e] * Pp + C[(2 * b + 1) * NE + e] * Pm;
                    Z[e] += Cz[(2 * b) * NE + e] * Qp + Cz[(2 * b + 1) * NE + e] * Qm;
                }
            }
            if (b + 1 < LZ) {
                Qp *= Pp;
                Pp *= Pp;
                Qm *= Pm;
                Pm *= Pm;
            }
        }
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            SA += Ex[e] * A[e];
            SAe += Exe[e] * A[e];
            SZ += Ex[e] * Z[e];
        }
    } else {
#pragma unroll
        for (int g = 0; g < NG; ++g)
            if (g < ng) {
                const R p1 = (R)1 + g_lam[g] * cs;
                const R pz1 = powu_(p1, g_iz[g] - 1);
                R K = 0, Ke = 0;
#pragma unroll
                for (int e = 0; e < NE; ++e) {
                    K += C[g * NE + e] * Ex[e];
                    Ke += C[g * NE + e] * Exe[e];
                }
                const R pz = pz1 * p1;
                SA += K * pz;
                SAe += Ke * pz;
                SZ += g_zl[g] * K * pz1;
            }
    }
}

// Sums over 8-lane groups of a wavefront for the tiled pair walk (every
// lane active): sum_lanes8 over lanes 8 g .. 8 g + 7 by exchanges on the
// vector unit, sum_stride8 over the lanes of one residue mod 8 (the
// exchange across rows of 16 lanes goes through the LDS crossbar); every
// lane of a group receives the sum.
template <int C>
__device__ inline float dpp_f(float v) {
    // every lane reads a lane of its own row, so no lane keeps an old value
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), C, 0xF, 0xF, true));
}
__device__ inline float sum_lanes8(float v) {
    v += dpp_f<0xB1>(v);   // quad_perm [1, 0, 3, 2]
    v += dpp_f<0x4E>(v);   // quad_perm [2, 3, 0, 1]
    v += dpp_f<0x141>(v);  // row_half_mirror: the other quad of the 8 lanes
    return v;
}
__device__ inline float sum_stride8(float v) {
    v += dpp_f<0x128>(v);  // row_ror:8: the lane 8 away in the row of 16
    v += __shfl_xor(v, 16, kWave);
    v += __shfl_xor(v, 32, kWave);
    return v;
}

// The force shares of one centre from its staged neighbours and its dE/dG,
// by one wavefront: wsrc(q) is element q of the centre's dE/dG (input
// order); shares go to Pout [i][slot], the centre's own to own [i], and
// with VIR the virial shares to PV and OV.
template <typename R, int NR, int NG, int NE, bool VIR, int LZ, bool K2, int ZM = -1, class WF>
__device__ inline void grad_wave(const BpDesc& P, const Stage<R>& S, int i, int nc, int nd, int cap, int lane,
                                 WF&& wsrc, R* __restrict__ Pout, R* __restrict__ own, R* __restrict__ PV,
                                 R* __restrict__ OV) {
    const int d0 = cap - nd;  // K2: the other kind's neighbours, [d0, cap)
    const int nr = P.nr, ng = P.ng, ne = P.ne;
    R eta_r[NR], g_lam[NG], g_zl[NG], eta_e[NE];
    int g_iz[NG];
#pragma unroll
    for (int r = 0; r < NR; ++r) eta_r[r] = r < nr ? (R)P.eta_r[r] : (R)0;
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        g_lam[g] = g < ng ? (R)P.g_lam[g] : (R)0;
        g_zl[g] = g < ng ? (R)P.g_zl[g] : (R)0;
        g_iz[g] = g < ng ? P.g_iz[g] : 1;
    }
#pragma unroll
    for (int e = 0; e < NE; ++e) eta_e[e] = e < ne ? (R)P.eta_e[e] : (R)0;
    R w[NR], w1[K2 ? NR : 1];
#pragma unroll
    for (int q = 0; q < NR; ++q) w[q] = q < nr ? wsrc(P.sr[0] + q) : (R)0;
    if constexpr (K2)
#pragma unroll
        for (int q = 0; q < NR; ++q) w1[q] = q < nr ? wsrc(P.sr[1] + q) : (R)0;
    // dE/dG of the G4 terms of angular block `key` times their coefficients
    // on the (group, eta) grid
    // (ladder layout: Cz, the same times 2^b and -2^b for group rows 2b and
    // 2b + 1, the zeta lam factors of rung b)
    R C[NG * NE], Cz[NG * NE];
    auto load_C = [&](int key) {
#pragma unroll
        for (int g = 0; g < NG; ++g)
#pragma unroll
            for (int e = 0; e < NE; ++e) {
                const int t = g * kBpMaxEta + e;
                const int q = (g < ng && e < ne) ? P.t_q[t] : -1;
                C[g * NE + e] = q >= 0 ? wsrc(P.sa[key] + q) * (R)P.t_coef[t] : (R)0;
                Cz[g * NE + e] = (R)((g & 1) ? -(1 << (g >> 1)) : (1 << (g >> 1))) * C[g * NE + e];
            }
    };
    load_C(0);
    const R Rs = (R)P.Rs, Rc2 = (R)P.Rc2, pirc = (R)P.pi_rc;
    // radial part of every slot, owned by one lane (K2: the other kind's
    // neighbours with their block's dE/dG)
    auto radial = [&](int a, const R* __restrict__ wr) {
        const R ax = S.x[a], ay = S.y[a], az = S.z[a], da = S.d[a], fa = S.fc[a], dfa = S.dfc[a];
        R sr = 0;
#pragma unroll
        for (int r = 0; r < NR; ++r)
            if (LZ > 0 || r < nr) {
                const R eta = eta_r[r];
                sr += wr[r] * dexp_(-eta * (da - Rs) * (da - Rs)) * ((R)-2 * eta * (da - Rs) * fa + dfa);
            }
        const R ia = S.id[a];
        S.g[a] = sr * ax * ia;
        S.g[cap + a] = sr * ay * ia;
        S.g[2 * cap + a] = sr * az * ia;
        if constexpr (VIR) {
            R* q = PV + ((size_t)i * cap + S.slot[a]) * 3;
            q[0] = sr * ax * ax * ia;
            q[1] = sr * ay * ay * ia;
            q[2] = sr * az * az * ia;
        }
    };
    for (int a = lane; a < nc; a += kWave) radial(a, w);
    if constexpr (K2)
        for (int a = d0 + lane; a < cap; a += kWave) radial(a, w1);
    __builtin_amdgcn_wave_barrier();
    // angular virial of the centre, per lane: for a pair (a, b) with
    // c = r_b - r_a, per axis (alpha + beta_a) x_a^2 + (alpha + beta_b) x_b^2
    // + (alpha + gam) x_c^2 + del x_a x_b (pattern_dG4_s_ contracted)
    R vx = 0, vy = 0, vz = 0;
    auto vir_pair = [&](R h, R ax_, R ay_, R az_, R bx_, R by_, R bz_, R cx_, R cy_, R cz_, R ia_, R ib_, R ic_,
                        R fa_, R fb_, R dfa_, R dfb_, R fcc_, R dfcc_, R cs_, R F_, R SA_, R SAe_, R SZ_) {
        const R al = (R)-2 * F_ * SAe_;
        const R ba = al - SZ_ * cs_ * ia_ * ia_ + SA_ * fb_ * fcc_ * dfa_ * ia_;
        const R bb = al - SZ_ * cs_ * ib_ * ib_ + SA_ * fa_ * fcc_ * dfb_ * ib_;
        const R gc = al + SA_ * fa_ * fb_ * dfcc_ * ic_;
        const R de = (R)2 * SZ_ * ia_ * ib_;
        vx += h * (ba * ax_ * ax_ + bb * bx_ * bx_ + gc * cx_ * cx_ + de * ax_ * bx_);
        vy += h * (ba * ay_ * ay_ + bb * by_ * by_ + gc * cy_ * cy_ + de * ay_ * by_);
        vz += h * (ba * az_ * az_ + bb * bz_ * bz_ + gc * cz_ * cz_ + de * az_ * bz_);
    };
    if constexpr (sizeof(R) == 4) {
    // angular part, FP32: every unordered pair {a, b} once, in tiles of 8
    // rows a by 8 columns b, one pair per lane (lane = 8 column + row); with
    // A_q = c_q pz_q, Z_q = c_q zeta lam pz1_q (c_q = w_q coef_q Ex_e(q))
    // summed over q,
    // d/dr_a = (-Z F cs / da - 2 F da SAe + fb dfa fcc SA) u_a
    //          + Z F / da u_b + (2 F dc SAe - fa fb dfcc SA) u_c, c = r_b - r_a,
    // and d/dr_b the same with a and b exchanged (u_c changes sign).  A
    // lane keeps its row a through a row of tiles, so the a-sides sum in
    // registers and meet across lanes once per row of tiles; the b-sides of
    // a tile sum over its 8 adjacent lanes of one column by lane exchanges
    // on the vector unit.  Nearly every tile holds a pair in range (of the
    // 55 tiles of 77 neighbours, about 52), so none is skipped, and all the
    // column values of a tile are read at once.
    // Rows a in [r0, r0 + nrow), columns b in [c0, c0 + ncol); same: the two
    // ranges are one and only pairs a < b are taken.
    struct GVals {
        R x, y, z, d, id, fc, dfc, E[NE];
    };
    auto walk = [&](int r0, int nrow, int c0, int ncol, bool same) {
    if (nrow <= 0 || ncol <= 0) return;
    const int ar = lane & 7, bc = lane >> 3;
    const int tr = (nrow + 7) >> 3, tc = (ncol + 7) >> 3;
    // column values of tile tj for this lane's column
    auto load_g = [&](int tj) {
        const int cb = tj * 8 + bc;
        const int b = c0 + (cb < ncol ? cb : 0);
        GVals v;
        v.x = S.x[b];
        v.y = S.y[b];
        v.z = S.z[b];
        v.d = S.d[b];
        v.id = S.id[b];
        v.fc = S.fc[b];
        v.dfc = S.dfc[b];
#pragma unroll
        for (int e = 0; e < NE; ++e) v.E[e] = (LZ > 0 || e < ne) ? S.E[e * cap + b] : (R)0;
        return v;
    };
    for (int ti = 0; ti < tr; ++ti) {
        const int ra = ti * 8 + ar;
        const bool va = ra < nrow;
        const int a = r0 + (va ? ra : 0);
        const R ax = S.x[a], ay = S.y[a], az = S.z[a], da = S.d[a], ia = S.id[a], fa = S.fc[a], dfa = S.dfc[a];
        R Ea[NE];
#pragma unroll
        for (int e = 0; e < NE; ++e) Ea[e] = (LZ > 0 || e < ne) ? S.E[e * cap + a] : (R)0;
        R hx = 0, hy = 0, hz = 0;
        for (int tj = same ? ti : 0; tj < tc; ++tj) {
            const GVals bv = load_g(tj);
            const int cb = tj * 8 + bc;
            const bool vb = cb < ncol;
            const int b = c0 + (vb ? cb : 0);
            const R bx = bv.x, by = bv.y, bz = bv.z, db = bv.d, ib = bv.id, fb = bv.fc, dfb = bv.dfc;
            const R cx = bx - ax, cy = by - ay, cz = bz - az;
            const R d2c = cx * cx + cy * cy + cz * cz;
            const bool ok = va && vb && (!same || ra < cb) && d2c >= (R)1e-15 && d2c < Rc2;
            const R ic = drsq_(d2c);
            const R dc = d2c * ic;
            const R fcc = (R)0.5 * (dcos_(pirc * dc) + (R)1);
            const R dfcc = (R)-0.5 * dsin_(pirc * dc) * pirc;
            const R cs = (ax * bx + ay * by + az * bz) * ia * ib;
            const R F = fa * fb * fcc;
            R Ex[NE], Exe[NE];
#pragma unroll
            for (int e = 0; e < NE; ++e) {
                Ex[e] = (LZ > 0 || e < ne) ? Ea[e] * bv.E[e] * dexp_(-eta_e[e] * d2c) : (R)0;
                Exe[e] = Ex[e] * eta_e[e];
            }
            R SA = 0, SAe = 0, SZ = 0;
            group_sums<R, NG, NE, LZ, ZM>(cs, C, Cz, Ex, Exe, g_lam, g_zl, g_iz, ng, SA, SAe, SZ);
            SZ *= F;
            const R cc = ((R)2 * F * dc * SAe - fa * fb * dfcc * SA) * ic;
            const R ka = (-SZ * cs * ia - (R)2 * F * da * SAe + fb * dfa * fcc * SA) * ia;
            const R kb = SZ * ia * ib;
            const R la = (-SZ * cs * ib - (R)2 * F * db * SAe + fa * dfb * fcc * SA) * ib;
            if constexpr (VIR)
                if (ok)
                    vir_pair((R)1, ax, ay, az, bx, by, bz, cx, cy, cz, ia, ib, ic, fa, fb, dfa, dfb, fcc, dfcc, cs, F,
                             SA, SAe, SZ);
            // with c = b - a: d/dr_a = (ka - cc) a + (kb + cc) b and
            // d/dr_b = (kb + cc) a + (la - cc) b
            const R pa = ok ? ka - cc : (R)0, pb = ok ? kb + cc : (R)0, qb = ok ? la - cc : (R)0;
            hx += pa * ax + pb * bx;
            hy += pa * ay + pb * by;
            hz += pa * az + pb * bz;
            const R gx = sum_lanes8(pb * ax + qb * bx);
            const R gy = sum_lanes8(pb * ay + qb * by);
            const R gz = sum_lanes8(pb * az + qb * bz);
            if (ar == 0 && vb) {
                atomicAdd(&S.g[b], gx);
                atomicAdd(&S.g[cap + b], gy);
                atomicAdd(&S.g[2 * cap + b], gz);
            }
        }
        hx = sum_stride8(hx);
        hy = sum_stride8(hy);
        hz = sum_stride8(hz);
        if (bc == 0 && va) {
            atomicAdd(&S.g[a], hx);
            atomicAdd(&S.g[cap + a], hy);
            atomicAdd(&S.g[2 * cap + a], hz);
        }
    }
    };
    walk(0, nc, 0, nc, true);
    if constexpr (K2) {
        load_C(1);
        walk(0, nc, d0, nd, false);
        load_C(2);
        walk(d0, nd, d0, nd, true);
    }
    } else {
    // angular part, FP64: every unordered pair {a, b} once; d/dr_b is d/dr_a
    // with a and b exchanged (u_c changes sign).  Pairs are walked by
    // diagonal, d = b - a, then by a, so the lanes of one step hold
    // consecutive a and consecutive b and the LDS accumulators take both
    // sides without two lanes hitting one address.
    auto pair64 = [&](int a, int b) {
        const R ax = S.x[a], ay = S.y[a], az = S.z[a], da = S.d[a];
        const R bx = S.x[b], by = S.y[b], bz = S.z[b], db = S.d[b];
        const R cx = bx - ax, cy = by - ay, cz = bz - az;
        const R d2c = cx * cx + cy * cy + cz * cz;
        if (!(d2c >= (R)1e-15 && d2c < Rc2)) return;
        const R fa = S.fc[a], dfa = S.dfc[a], fb = S.fc[b], dfb = S.dfc[b];
        const R ia = S.id[a], ib = S.id[b], ic = drsq_(d2c);
        const R dc = d2c * ic;
        const R fcc = (R)0.5 * (dcos_(pirc * dc) + (R)1);
        const R dfcc = (R)-0.5 * dsin_(pirc * dc) * pirc;
        const R cs = (ax * bx + ay * by + az * bz) * ia * ib;
        const R F = fa * fb * fcc;
        R Ex[NE], Exe[NE];
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            Ex[e] = (LZ > 0 || e < ne) ? S.E[e * cap + a] * S.E[e * cap + b] * dexp_(-eta_e[e] * d2c) : (R)0;
            Exe[e] = Ex[e] * eta_e[e];
        }
        R SA = 0, SAe = 0, SZ = 0;
        group_sums<R, NG, NE, LZ, ZM>(cs, C, Cz, Ex, Exe, g_lam, g_zl, g_iz, ng, SA, SAe, SZ);
        SZ *= F;
        const R cc = ((R)2 * F * dc * SAe - fa * fb * dfcc * SA) * ic;
        const R ka = (-SZ * cs * ia - (R)2 * F * da * SAe + fb * dfa * fcc * SA) * ia;
        const R kb = SZ * ia * ib;
        const R la = (-SZ * cs * ib - (R)2 * F * db * SAe + fa * dfb * fcc * SA) * ib;
        if constexpr (VIR)
            vir_pair((R)1, ax, ay, az, bx, by, bz, cx, cy, cz, ia, ib, ic, fa, fb, dfa, dfb, fcc, dfcc, cs, F, SA, SAe,
                     SZ);
        atomicAdd(&S.g[a], ka * ax + kb * bx + cc * cx);
        atomicAdd(&S.g[cap + a], ka * ay + kb * by + cc * cy);
        atomicAdd(&S.g[2 * cap + a], ka * az + kb * bz + cc * cz);
        atomicAdd(&S.g[b], la * bx + kb * ax - cc * cx);
        atomicAdd(&S.g[cap + b], la * by + kb * ay - cc * cy);
        atomicAdd(&S.g[2 * cap + b], la * bz + kb * az - cc * cz);
    };
    // a triangle [lo, lo + n) by diagonal; a rectangle [la, la + na) x
    // [lb, lb + nb) by shifted rows (b = (a + d) mod nb), so that the lanes
    // of one step hold different a and, for na <= nb, different b
    auto tri = [&](int lo, int n) {
        for (PairWalk pw(n, lane); pw.valid(); pw.advance()) pair64(lo + pw.t, lo + pw.t + pw.a + 1);
    };
    auto rect = [&](int la, int na, int lb, int nb) {
        for (int t = lane; t < na * nb; t += kWave) {
            const int d = t / na, a = t - d * na;
            pair64(la + a, lb + (a + d) % nb);
        }
    };
    tri(0, nc);
    if constexpr (K2) {
        load_C(1);
        rect(0, nc, d0, nd);
        load_C(2);
        tri(d0, nd);
    }
    }
    __builtin_amdgcn_wave_barrier();
    R ox = 0, oy = 0, oz = 0;
    for (int a = lane; a < (K2 ? cap : nc); a += kWave) {
        if (K2 && a >= nc && a < d0) continue;
        const R gx = S.g[a], gy = S.g[cap + a], gz = S.g[2 * cap + a];
        R* q = Pout + ((size_t)i * cap + S.slot[a]) * 3;
        q[0] = -gx;
        q[1] = -gy;
        q[2] = -gz;
        ox += gx;
        oy += gy;
        oz += gz;
    }
    ox = wave_sum(ox);
    oy = wave_sum(oy);
    oz = wave_sum(oz);
    if (lane == 0) {
        own[(size_t)i * 3 + 0] = ox;
        own[(size_t)i * 3 + 1] = oy;
        own[(size_t)i * 3 + 2] = oz;
    }
    if constexpr (VIR) {
        vx = wave_sum(vx);
        vy = wave_sum(vy);
        vz = wave_sum(vz);
        if (lane == 0) {
            OV[(size_t)i * 3 + 0] = vx;
            OV[(size_t)i * 3 + 1] = vy;
            OV[(size_t)i * 3 + 2] = vz;
        }
    }
}

template <typename R, int NR, int NG, int NE, bool VIR, int LZ = 0, bool K2 = false>
__global__ void __launch_bounds__(256) k_bp_grad(const double4* __restrict__ pos, const int* __restrict__ nl,
                                                 const int* __restrict__ nn, int tpa, int M, int cap,
                                                 const SysDesc* __restrict__ desc, const BpDesc* __restrict__ Pd,
                                                 const R* __restrict__ wgt, int ldg, R* __restrict__ Pout,
                                                 R* __restrict__ own, R* __restrict__ PV, R* __restrict__ OV, int n,
                                                 StepArgs SA, int* __restrict__ misc) {
    if (skip_step(SA, misc)) return;
    extern __shared__ char smem[];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + wv;
    if (i >= n) return;
    const BpDesc& P = *Pd;
    const Stage<R> S = carve<R>(smem + wv * stage_bytes(cap, P.ne, sizeof(R), true), cap, P.ne, true);
    const double4 pi = pos[i];
    const SysDesc& s = desc[w_sys(pi.w)];
    const int cnt = min(nn[i], cap);
    int nd = 0;
    const int nc = bp_stage<R>(i, pi, s, nl, cnt, tpa, M, pos, P, S, Pout, VIR ? PV : (R*)nullptr, cap,
                               K2 ? (int)w_kind(pi.w) : -1, nd);
    grad_wave<R, NR, NG, NE, VIR, LZ, K2>(P, S, i, nc, nd, cap, lane,
                                          [&](int q) { return wgt[(size_t)i * ldg + q]; }, Pout, own, PV, OV);
}

// lane p's value of v, p uniform over the wavefront
__device__ inline float readlane_f(float v, int p) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), p));
}

// s + sum over p < cnt of w(p) times lane p's v, in order of p: the loop
// runs to the widest layer in chunks of 8 (a loop with a run-time count and
// the lane reads would not be unrolled), and each chunk's 8 weight reads are
// issued before its products, so they are in flight together
template <class W>
__device__ inline float lane_dot(float s, float v, int cnt, W&& w) {
#pragma unroll
    for (int p0 = 0; p0 < kBpMaxHidden; p0 += 8) {
        if (p0 < cnt) {
            float wv[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) wv[j] = p0 + j < cnt ? w(p0 + j) : 0.f;
#pragma unroll
            for (int j = 0; j < 8; ++j) s = fmaf(wv[j], readlane_f(v, p0 + j), s);
        }
    }
    return s;
}

// The network of one centre by one wavefront in FP32 (maise_ann.cxx:8-55,
// maise_bpnn.cxx:275-325): lane j holds unit j of each layer (widths up to
// 64) and takes the previous layer's outputs one at a time from the lane
// that holds them.  The input vector comes from the descriptor sums acc
// (every lane holds them) in input order: radial term r at r, angular grid
// term t at nr + t_q[t] with its coefficient.  w: the FP64 block's layout
// in FP32 (W01 [h1][nin] range-scaled, b01, W12 [h2][h1], b12, W23, b23,
// W01T [nin][h1], W12T [h1][h2]).  E = b23 + W23 . h2; returns dE/dx_k in
// lane k (0 in lanes k >= nin).
template <int NR, int NG, int NE>
__device__ inline float mlp_wave(const BpDesc& P, const float* __restrict__ w, const float (&acc)[NR + NG * NE],
                                 int lane, float& E) {
    const int nin = P.nin, h1 = P.h1, h2 = P.h2, nr = P.nr;
    const float* W01 = w;
    const float* b01 = W01 + h1 * nin;
    const float* W12 = b01 + h1;
    const float* b12 = W12 + h2 * h1;
    const float* W23 = b12 + h2;
    const float b23 = W23[h2];
    const float* W01T = W23 + h2 + 1;
    const float* W12T = W01T + nin * h1;
    // lanes past a layer's width read its last unit and discard the result
    const int j1 = min(lane, h1 - 1), j2 = min(lane, h2 - 1), k = min(lane, nin - 1);
    float s1 = b01[j1];
#pragma unroll
    for (int r = 0; r < NR; ++r)
        if (r < nr) s1 = fmaf(W01T[r * h1 + j1], acc[r], s1);
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const int t = g * kBpMaxEta + e;
            const int q = (g < P.ng && e < P.ne) ? P.t_q[t] : -1;
            if (q >= 0) s1 = fmaf(W01T[(nr + q) * h1 + j1], (float)P.t_coef[t] * acc[NR + g * NE + e], s1);
        }
    const float y1 = lane < h1 ? tanhf(s1) : 0.f;
    const float s2 = lane_dot(b12[j2], y1, h1, [&](int p) { return W12T[p * h2 + j2]; });
    const float y2 = lane < h2 ? tanhf(s2) : 0.f;
    const float e2 = lane < h2 ? W23[j2] * y2 : 0.f;
    const float t2 = lane < h2 ? W23[j2] * (1.f - y2 * y2) : 0.f;
    E = b23 + wave_sum(e2);
    const float s3 = lane_dot(0.f, t2, h2, [&](int q) { return W12[q * h1 + j1]; });
    const float d1 = lane < h1 ? s3 * (1.f - y1 * y1) : 0.f;
    const float so = lane_dot(0.f, d1, h1, [&](int p) { return W01[p * nin + k]; });
    return lane < nin ? so : 0.f;
}

// passes 1-3 in one kernel (FP32, one species): one wavefront per centre
// stages the neighbours once, forms the descriptors, runs the network with
// dE/dG kept in registers and writes the force shares, as k_bp_desc,
// k_bp_mlp_mfma and k_bp_grad do one after the other (maise_bpnn.cxx:51-96,
// 122-325).  G and dE/dG are written as well, for inspection.
template <int NR, int NG, int NE, bool VIR, int LZ, int ZM = -1>
__global__ void __launch_bounds__(256) k_bp_fused(const double4* __restrict__ pos, const int* __restrict__ nl,
                                                  const int* __restrict__ nn, int tpa, int M, int cap,
                                                  const SysDesc* __restrict__ desc, const BpDesc* __restrict__ Pd,
                                                  const float* __restrict__ mw, float* __restrict__ G,
                                                  float* __restrict__ wout, int ldg, double* __restrict__ Ep,
                                                  float* __restrict__ Pout, float* __restrict__ own,
                                                  float* __restrict__ PV, float* __restrict__ OV, int n, StepArgs SA,
                                                  int* __restrict__ misc) {
    if (skip_step(SA, misc)) return;
    extern __shared__ char smem[];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + wv;
    if (i >= n) return;
    const BpDesc& P = *Pd;
    const Stage<float> S = carve<float>(smem + wv * stage_bytes(cap, P.ne, sizeof(float), true), cap, P.ne, true);
    const double4 pi = pos[i];
    const SysDesc& s = desc[w_sys(pi.w)];
    const int cnt = min(nn[i], cap);
    int nd = 0;
    const int nc = bp_stage<float>(i, pi, s, nl, cnt, tpa, M, pos, P, S, Pout, VIR ? PV : (float*)nullptr, cap, -1,
                                   nd);
    float acc[NR + NG * NE];
    desc_wave<float, NR, NG, NE, LZ, false, ZM>(P, S, nc, nd, cap, lane, G + (size_t)i * ldg, ldg, acc);
    // inactive centres: no energy and zero dE/dG
    const bool act = !w_inactive(pi.w);
    float E = 0.f;
    const float dx = act ? mlp_wave<NR, NG, NE>(P, mw, acc, lane, E) * (float)P.escale : 0.f;
    if (lane < ldg) wout[(size_t)i * ldg + lane] = dx;
    if (act && lane == 0) Ep[i] = (double)E * P.escale;
    grad_wave<float, NR, NG, NE, VIR, LZ, false, ZM>(P, S, i, nc, nd, cap, lane,
                                                     [&](int q) { return readlane_f(dx, q); }, Pout, own, PV, OV);
}

// reverse slot index: rev(m, k) = slot of m in the list of its k-th
// neighbour i (-1 when a capacity cut left m out of i's list).  One thread
// per (m, k) searches i's list, four slots per step so the reads overlap.
// With four or more lanes per atom, slots q0..q0+3 (q0 a multiple of 4) of
// one atom are adjacent and 16-byte aligned in the wavefront-packed layout
// (list_index; the capacity is a multiple of the lanes per atom, so they
// are all allocated), and one 16-byte load reads them: a wave's search
// step is then one memory instruction instead of four to the same lines.
__global__ void k_bp_rev(const int* __restrict__ nl, const int* __restrict__ nn, int* __restrict__ rev, int tpa, int M,
                         int cap, int n) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int m = (int)(t / cap), k = (int)(t % cap);
    if (m >= n) return;
    const int cm = min(nn[m], cap);
    if (k >= cm) return;
    const long long lk = list_index(m, k, tpa, M);
    const int i = ent_j(nl[lk]);
    const int ci = min(nn[i], cap);
    int r = -1;
    if (tpa >= 8) {
        // eight adjacent slots (q0 a multiple of 8), two 16-byte loads
        for (int q0 = 0; q0 < ci && r < 0; q0 += 8) {
            const int4* p = reinterpret_cast<const int4*>(nl + list_index(i, q0, tpa, M));
            const int4 v0 = p[0], v1 = p[1];
            const int e[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
            for (int u = 7; u >= 0; --u)
                if (q0 + u < ci && ent_j(e[u]) == m) r = q0 + u;
        }
        rev[lk] = r;
        return;
    }
    if (tpa >= 4) {
        for (int q0 = 0; q0 < ci && r < 0; q0 += 4) {
            const int4 v = *reinterpret_cast<const int4*>(nl + list_index(i, q0, tpa, M));
            const int e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int u = 3; u >= 0; --u)
                if (q0 + u < ci && ent_j(e[u]) == m) r = q0 + u;
        }
        rev[lk] = r;
        return;
    }
    for (int q0 = 0; q0 < ci && r < 0; q0 += 4) {
        int v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = q0 + u < ci ? ent_j(nl[list_index(i, q0 + u, tpa, M)]) : -1;
#pragma unroll
        for (int u = 3; u >= 0; --u)
            if (v[u] == m) r = q0 + u;
    }
    rev[lk] = r;
}

// pass 4: gather the force shares, then the fused velocity-Verlet update.
// VIR: also the virial shares; vmode 1 writes V (its diagonal from the
// shares, the rest zero, as the reference's prepare leaves it) with the
// force, vmode 2 adds the diagonal to V and leaves the force (the virials
// call of solver::Ternary, ternary.cxx:57-67).
template <typename R, int TPA, bool VIR>
__global__ void __launch_bounds__(256) k_bp_gather(const double4* __restrict__ pos, double4* __restrict__ pos_next,
                                                   const int* __restrict__ nl, const int* __restrict__ nn,
                                                   const int* __restrict__ rev, int M, int cap,
                                                   const SysDesc* __restrict__ desc, const R* __restrict__ Pin,
                                                   const R* __restrict__ own, const R* __restrict__ PV,
                                                   const R* __restrict__ OV, int vmode, AtomArrays A, StepArgs SA,
                                                   const double4* __restrict__ rsave, int* __restrict__ misc) {
    if (skip_step(SA, misc)) return;
    // TPA lanes per atom walk its list slots in the wavefront-packed layout
    // (coalesced list and reverse-index reads), then sum their shares
    const AtomSlots as = atom_slots<TPA>(M);
    const int m = as.i;
    const bool valid = m < A.live;
    const double4 pm = valid ? pos[m] : make_double4(0, 0, 0, 0);
    const bool active = valid && !w_inactive(pm.w);
    const int cm = active ? min(nn[m], cap) : 0;
    R fx = 0, fy = 0, fz = 0, vx = 0, vy = 0, vz = 0;
    for (int c = 0; c * TPA + as.sub < cm; ++c) {
        const long long li = as.base + (long long)c * kWave;
        const int i = ent_j(nl[li]);
        const int r = rev[li];
        if (r < 0) continue;
        const size_t o = ((size_t)i * cap + r) * 3;
        fx += Pin[o];
        fy += Pin[o + 1];
        fz += Pin[o + 2];
        if constexpr (VIR) {
            vx += PV[o];
            vy += PV[o + 1];
            vz += PV[o + 2];
        }
    }
    fx = wave_sum_group(fx, TPA);
    fy = wave_sum_group(fy, TPA);
    fz = wave_sum_group(fz, TPA);
    if constexpr (VIR) {
        vx = wave_sum_group(vx, TPA);
        vy = wave_sum_group(vy, TPA);
        vz = wave_sum_group(vz, TPA);
    }
    if (!valid || as.sub != 0) return;
    if (!active) {
        if (SA.mode == MODE_FINAL_DRIFT) pos_next[m] = pm;
        return;
    }
    if constexpr (VIR) {
        const int n = A.n;
        const double V[3] = {(double)vx + (double)OV[(size_t)m * 3 + 0], (double)vy + (double)OV[(size_t)m * 3 + 1],
                             (double)vz + (double)OV[(size_t)m * 3 + 2]};
        if (vmode == 2) {
            for (int c = 0; c < 3; ++c) A.d[(size_t)(D_VXX + c) * n + m] += V[c];
            return;
        }
        for (int c = 0; c < 3; ++c) A.d[(size_t)(D_VXX + c) * n + m] = V[c];
        for (int f = D_VXY; f <= D_VYZ; ++f) A.d[(size_t)f * n + m] = 0.0;
    }
    fx += own[(size_t)m * 3 + 0];
    fy += own[(size_t)m * 3 + 1];
    fz += own[(size_t)m * 3 + 2];
    const FinIn fin = load_fin(m, A, SA, rsave, true);
    const SysDesc& s = desc[w_sys(pm.w)];
    finalize_atom<R>(m, pm, (double)fx, (double)fy, (double)fz, A, SA, pos_next, fin, s, misc);
}

// --------------------------------------------------------------------------
// host: MAISE model file (parser/maise_model.cxx:29-129 layout) and setup
// --------------------------------------------------------------------------
namespace {

std::vector<std::string> words(const std::string& s) {
    std::vector<std::string> out;
    std::istringstream is(s);
    std::string t;
    while (is >> t) out.push_back(t);
    return out;
}

struct Maise {
    int nspc = 0;
    std::vector<int> NN;
    int nw = 0;
    std::vector<double> range;  // [nspc][inputs]
    std::vector<double> weights;
    double b2a = 1, scaling = 1;
    std::vector<std::vector<int>> map;
    std::vector<std::vector<double>> vals;
};

Maise read_maise(const char* path) {
    std::ifstream f(path);
    if (!f.is_open()) throw Error(E_FILE, std::string("Failed to open file: ") + path);
    Maise m;
    std::string line;
    auto next = [&]() {
        if (!std::getline(f, line)) throw Error(E_FILE, std::string("unexpected end of MAISE model file: ") + path);
    };
    auto field = [&]() {
        next();
        size_t p = line.find('|', 1);
        if (p == std::string::npos) throw Error(E_FILE, "malformed MAISE head line: " + line);
        return line.substr(p + 1, p + 10);
    };
    for (int i = 0; i < 3; ++i) next();
    (void)field();  // model id
    m.nspc = std::stoi(field());
    (void)field();  // species types
    (void)field();  // species names
    (void)field();  // number of layers
    for (auto& t : words(field())) m.NN.push_back(std::stoi(t));
    m.nw = std::stoi(field());
    const std::string ICR = "|                          input component range                            |";
    do next(); while (line != ICR);
    next();
    next();
    if (m.NN.empty()) throw Error(E_FILE, "MAISE architecture missing");
    for (int s = 0; s < m.nspc; ++s)
        for (int k = 0; k < m.NN[0]; ++k) {
            next();
            auto t = words(line);
            if (t.size() < 3) throw Error(E_FILE, "malformed MAISE input range row");
            m.range.push_back(std::stod(t[2]));
        }
    for (int i = 0; i < 3; ++i) next();
    for (int i = 0; i < m.nw; ++i) {
        next();
        m.weights.push_back(std::stod(line));
    }
    for (int i = 0; i < 3; ++i) next();
    auto bs = words(line);
    if (bs.size() < 2) throw Error(E_FILE, "malformed MAISE B2A/Scaling row");
    m.b2a = std::stod(bs[0]);
    m.scaling = std::stod(bs[1]);
    for (int i = 0; i < 3; ++i) next();
    next();
    while (!line.empty()) {
        std::vector<int> row;
        for (auto& t : words(line)) row.push_back(std::stoi(t));
        m.map.push_back(row);
        next();
    }
    for (int i = 0; i < 7; ++i) {
        next();
        auto t = words(line);
        if (t.size() < 2 || std::stoi(t[1]) != (int)t.size() - 2)
            throw Error(E_FILE, "malformed MAISE descriptor value row: " + line);
        std::vector<double> v;
        for (size_t k = 2; k < t.size(); ++k) v.push_back(std::stod(t[k]));
        m.vals.push_back(v);
    }
    return m;
}

template <class T>
T* dnew(size_t count) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess)
        throw Error(E_NOMEM, "out of device memory for the BPNN model");
    return static_cast<T*>(p);
}

BpWork& work(Batch& b) {
    if (!b.bp_scratch) {
        b.bp_scratch = new BpWork();
    }
    return *static_cast<BpWork*>(b.bp_scratch);
}

void free_work(BpWork& w) {
    for (void* p : {w.a.G, w.a.w, w.a.P, w.a.own, w.a.PV, w.a.OV, (void*)w.a.rev})
        if (p) (void)hipFree(p);
    w.a = BpBufs{};
    w.cap = w.n = 0;
    w.prec = -1;
    w.rev_elems = 0;
    w.has_vir = false;
    w.rev_valid = false;
}

}  // namespace

void bpnn_release_work(Batch& b) {
    if (!b.bp_scratch) return;
    BpWork* w = static_cast<BpWork*>(b.bp_scratch);
    free_work(*w);
    delete w;
    b.bp_scratch = nullptr;
}

// The model of a MAISE file checked against the engine's limits, with the
// network weights in the layouts the kernels read (host side only).
static std::unique_ptr<BpnnParams> parse_model(const char* path, std::vector<double>& W64, std::vector<float>& W32,
                                               std::vector<half4_t>& B) {
    Maise m = read_maise(path);
    if (m.NN.size() != 4) throw Error(E_FILE, "MaiseBPNN: only 4-layer networks are supported");
    if (m.nspc < 1 || m.nspc > 2)
        throw Error(E_INVAL, "MaiseBPNN: one or two species are supported (with three or more the reference's "
                             "angular block of a pair depends on the order of the neighbour list)");
    if (m.NN[3] != 1) throw Error(E_FILE, "MaiseBPNN: the output layer must have one neuron");
    auto owner = std::make_unique<BpnnParams>();
    BpnnParams* bp = owner.get();
    BpDesc& D = bp->h;
    std::memset(&D, 0, sizeof(D));
    const int nin = m.NN[0], h1 = m.NN[1], h2 = m.NN[2];
    if (h1 > kBpMaxHidden || h2 > kBpMaxHidden || nin > kBpMaxIn)
        throw Error(E_INVAL, "MaiseBPNN: at most 64 inputs and 64 neurons per hidden layer are supported");
    // maise_model.cxx:34-50 unit scaling
    const double lin = m.b2a * m.scaling * 0.1, isq = 1.0 / lin / lin;
    if (m.vals.size() < 7) throw Error(E_FILE, "MAISE descriptor table incomplete");
    D.Rc = m.vals[0].at(0) * lin;
    D.Rs = m.vals[1].at(0) * lin;
    std::vector<double> eta1 = m.vals[2], eta2 = m.vals[3];
    for (auto& e : eta1) e *= isq;
    for (auto& e : eta2) e *= isq;
    std::vector<double> etas;
    for (auto& row : m.map) {
        if (row.size() != 9) throw Error(E_FILE, "MAISE descriptor map rows need 9 columns");
        if (row[1] == 2) {
            if (D.nr >= kBpMaxR) throw Error(E_INVAL, "at most 16 radial descriptors are supported");
            D.eta_r[D.nr++] = eta1.at(row[4]);
        } else if (row[1] == 4) {
            if (D.na >= kBpMaxA) throw Error(E_INVAL, "at most 64 angular descriptors are supported");
            double eta = eta2.at(row[5]), z = m.vals[5].at(row[7]), l = m.vals[6].at(row[8]);
            int e = -1;
            for (size_t q = 0; q < etas.size(); ++q)
                if (etas[q] == eta) e = (int)q;
            if (e < 0) {
                if ((int)etas.size() >= kBpMaxEta) throw Error(E_INVAL, "at most 8 distinct G4 eta values are supported");
                etas.push_back(eta);
                e = (int)etas.size() - 1;
            }
            const int a = D.na++;
            D.a_e[a] = e;
            D.a_zeta[a] = z;
            D.a_lam[a] = l;
            D.a_coef[a] = std::pow(2.0, 1.0 - z) * 2.0;  // base.cxx:120 factor 2
            if (!(z == std::floor(z) && z >= 1 && z <= 32))
                throw Error(E_INVAL, "G4 zeta must be a whole number from 1 to 32");
            D.a_iz[a] = (int)z;
        } else {
            throw Error(E_INVAL, "only G2 radial and G4 angular descriptors are supported");
        }
    }
    // nspc radial and nspc (nspc + 1) / 2 angular blocks
    D.nspc = m.nspc;
    const int blocks = D.nspc * D.nr + D.nspc * (D.nspc + 1) / 2 * D.na;
    if (blocks < nin) throw Error(E_FILE, "fewer descriptors than network inputs");
    if (blocks != nin) throw Error(E_INVAL, "descriptor count must equal the network input width");
    D.sr[0] = 0;
    D.sr[1] = D.nr + D.na;
    D.sa[0] = D.nr;
    D.sa[1] = 2 * D.nr + D.na;
    D.sa[2] = 2 * D.nr + 2 * D.na;
    D.ne = (int)etas.size();
    for (int e = 0; e < D.ne; ++e) D.eta_e[e] = etas[e];
    // G4 terms by (lambda, zeta) group and eta
    D.ng = 0;
    for (int t = 0; t < kBpGridG * kBpMaxEta; ++t) {
        D.t_q[t] = -1;
        D.t_coef[t] = 0.0;
    }
    // the ladder layout when every term allows it
    int zb_max = -1;
    bool ladder = D.na > 0;
    auto zeta_bit = [](int iz) {
        for (int b = 0; b < kBpMaxLadder; ++b)
            if (iz == (1 << b)) return b;
        return -1;
    };
    for (int q = 0; q < D.na; ++q) {
        const int b = zeta_bit(D.a_iz[q]);
        ladder = ladder && b >= 0 && (D.a_lam[q] == 1.0 || D.a_lam[q] == -1.0);
        zb_max = std::max(zb_max, b);
    }
    D.lz = ladder ? ladder_rungs(D.nr, D.ne, zb_max + 1) : 0;
    ladder = D.lz > 0;
    if (ladder) {
        D.ng = 2 * D.lz;
        for (int g = 0; g < D.ng; ++g) {
            D.g_lam[g] = (g & 1) ? -1.0 : 1.0;
            D.g_iz[g] = 1 << (g >> 1);
            D.g_zeta[g] = D.g_iz[g];
            D.g_zl[g] = D.g_zeta[g] * D.g_lam[g];
        }
    }
    for (int q = 0; q < D.na; ++q) {
        int g = -1;
        if (ladder) {
            g = 2 * zeta_bit(D.a_iz[q]) + (D.a_lam[q] < 0 ? 1 : 0);
        } else {
            for (int h = 0; h < D.ng; ++h)
                if (D.g_lam[h] == D.a_lam[q] && D.g_zeta[h] == D.a_zeta[q]) g = h;
        }
        if (g < 0) {
            if (D.ng >= kBpMaxG) throw Error(E_INVAL, "at most 8 distinct (lambda, zeta) pairs among the G4 terms");
            g = D.ng++;
            D.g_lam[g] = D.a_lam[q];
            D.g_zeta[g] = D.a_zeta[q];
            D.g_zl[g] = D.a_zeta[q] * D.a_lam[q];
            D.g_iz[g] = D.a_iz[q];
        }
        const int t = g * kBpMaxEta + D.a_e[q];
        if (D.t_q[t] >= 0) throw Error(E_INVAL, "two G4 terms with the same eta, lambda and zeta");
        D.t_q[t] = q;
        D.t_coef[t] = D.a_coef[q];
    }
    D.Rc2 = D.Rc * D.Rc;
    D.pi_rc = M_PI / D.Rc;
    D.nin = nin;
    D.h1 = h1;
    D.h2 = h2;
    D.escale = 10.0 * 1.602176634 * 6.02214076;  // maise_bpnn.cxx:317
    // maise_ann.cxx:23-51 weight order (species after species), columns of
    // W01 scaled by the species' range; each species' arrays below are one
    // block of the same size, species after species
    size_t w = 0;
    auto take = [&](size_t k) {
        if (w + k > m.weights.size()) throw Error(E_FILE, "MAISE weight count does not match the architecture");
        std::vector<double> v(m.weights.begin() + w, m.weights.begin() + w + k);
        w += k;
        return v;
    };
    auto pad16 = [](int x) { return (x + 15) / 16 * 16; };
    bp->nin_pad = pad16(nin);
    bp->h1_pad = pad16(h1);
    bp->h2_pad = pad16(h2);
    W64.clear();
    W32.clear();
    B.clear();
    for (int sp = 0; sp < m.nspc; ++sp) {
        auto b01 = take(h1), b12 = take(h2), b23 = take(1);
        auto W01 = take((size_t)h1 * nin), W12 = take((size_t)h2 * h1), W23 = take(h2);
        const double* range = &m.range.at((size_t)sp * nin);
        for (int k = 0; k < nin; ++k)
            for (int q = 0; q < h1; ++q) W01[(size_t)q * nin + k] *= range[k];
        const size_t w64_0 = W64.size(), w32_0 = W32.size(), b_0 = B.size();
        for (auto* v : {&W01, &b01, &W12, &b12, &W23}) W64.insert(W64.end(), v->begin(), v->end());
        W64.push_back(b23[0]);
        // transposed copies for the FP64 network's column reads
        for (int k = 0; k < nin; ++k)
            for (int q = 0; q < h1; ++q) W64.push_back(W01[(size_t)q * nin + k]);
        for (int p = 0; p < h1; ++p)
            for (int q = 0; q < h2; ++q) W64.push_back(W12[(size_t)q * h1 + p]);
        // The MFMA network takes input k as G_k xs_k with xs_k the power of
        // two nearest the file's range_k, and W01 column k divided by it
        // (both exact): its inputs are then of order one, as the range
        // normalises them, and the FP16 low halves of small descriptors stay
        // normal.
        std::vector<double> xs(bp->nin_pad, 1.0);
        for (int k = 0; k < nin; ++k) {
            const double r = range[k];
            xs[k] = (r > 0 && std::isfinite(r)) ? std::ldexp(1.0, (int)std::lround(std::log2(r))) : 1.0;
        }
        for (auto* v : {&b01, &b12, &W23}) W32.insert(W32.end(), v->begin(), v->end());
        W32.push_back((float)b23[0]);
        for (double x : xs) W32.push_back((float)x);
        struct Gm {
            int K, N;
            std::function<double(int, int)> at;
        };
        std::vector<Gm> gm = {
            {nin, h1, [&](int k, int n) { return W01[(size_t)n * nin + k] / xs[k]; }},  // X W01^T
            {h1, h2, [&](int k, int n) { return W12[(size_t)n * h1 + k]; }},            // H1 W12^T
            {h2, h1, [&](int k, int n) { return W12[(size_t)k * h1 + n]; }},            // D2 W12
            {h1, nin, [&](int k, int n) { return W01[(size_t)k * nin + n] / xs[n]; }},  // D1 W01
        };
        // B operands of the four GEMMs in the MFMA lane layout, split hi/lo
        for (int g = 0; g < 4; ++g) {
            const int kt = pad16(gm[g].K) / 16, nt = pad16(gm[g].N) / 16;
            bp->boff[g] = (int)(B.size() - b_0);
            bp->kt[g] = kt;
            bp->nt[g] = nt;
            for (int t = 0; t < nt; ++t)
                for (int k = 0; k < kt; ++k)
                    for (int part = 0; part < 2; ++part)
                        for (int lane = 0; lane < 64; ++lane) {
                            half4_t v;
                            for (int c = 0; c < 4; ++c) {
                                const int kk = k * 16 + (lane >> 4) * 4 + c, nn = t * 16 + (lane & 15);
                                const double x = (kk < gm[g].K && nn < gm[g].N) ? gm[g].at(kk, nn) : 0.0;
                                const _Float16 hi = (_Float16)(float)x;
                                const _Float16 lo = (_Float16)((float)x - (float)hi);
                                v[c] = part == 0 ? hi : lo;
                            }
                            B.push_back(v);
                        }
        }
        bp->w64_stride = W64.size() - w64_0;
        bp->w32_stride = W32.size() - w32_0;
        bp->b_stride = B.size() - b_0;
    }
    if ((int)w != m.nw) throw Error(E_FILE, "MAISE weight count does not match the architecture");
    return owner;
}

Pot* bpnn_create(Ctx* ctx, const char* path) {
    std::vector<double> W64;
    std::vector<float> W32;
    std::vector<half4_t> B;
    std::unique_ptr<BpnnParams> owner = parse_model(path, W64, W32, B);
    BpnnParams* bp = owner.get();
    const BpDesc& D = bp->h;
    auto* p = new Pot();
    p->ctx = ctx;
    p->type = PT_BPNN;
    p->rcut = D.Rc;
    p->bp = owner.release();
    try {
        bp->d = dnew<BpDesc>(1);
        bp->w64 = dnew<double>(W64.size());
        bp->w32 = dnew<float>(W32.size());
        bp->w64f = dnew<float>(W64.size());
        bp->bsplit = dnew<half4_t>(B.size());
        HIPCHECK(hipMemcpy(bp->d, &D, sizeof(D), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(bp->w64, W64.data(), W64.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(bp->w32, W32.data(), W32.size() * sizeof(float), hipMemcpyHostToDevice));
        const std::vector<float> W64f(W64.begin(), W64.end());
        HIPCHECK(hipMemcpy(bp->w64f, W64f.data(), W64f.size() * sizeof(float), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(bp->bsplit, B.data(), B.size() * sizeof(half4_t), hipMemcpyHostToDevice));
    } catch (...) {
        bpnn_destroy(p);
        throw;
    }
    return p;
}

void bpnn_describe(const char* path, int* arch, int* counts, double* rcut) {
    std::vector<double> W64;
    std::vector<float> W32;
    std::vector<half4_t> B;
    std::unique_ptr<BpnnParams> bp = parse_model(path, W64, W32, B);
    const BpDesc& D = bp->h;
    if (arch) {
        arch[0] = D.nin;
        arch[1] = D.h1;
        arch[2] = D.h2;
        arch[3] = 1;
    }
    if (counts) {
        counts[0] = D.nr;
        counts[1] = D.na;
        counts[2] = D.ne;
        // distinct (lambda, zeta) pairs, whatever the kernels' group layout
        int distinct = 0;
        for (int q = 0; q < D.na; ++q) {
            bool seen = false;
            for (int r = 0; r < q; ++r) seen = seen || (D.a_lam[r] == D.a_lam[q] && D.a_zeta[r] == D.a_zeta[q]);
            distinct += seen ? 0 : 1;
        }
        counts[3] = distinct;
    }
    if (rcut) *rcut = D.Rc;
}

int bpnn_species(const Pot& p) { return p.bp ? p.bp->h.nspc : 1; }

void bpnn_destroy(Pot* p) {
    if (p->bp) {
        BpnnParams* bp = p->bp;
        for (void* q : {(void*)bp->d, (void*)bp->w64, (void*)bp->w32, (void*)bp->w64f, (void*)bp->bsplit})
            if (q) (void)hipFree(q);
        delete bp;
    }
    delete p;
}

static void ensure_work(Batch& b, const BpnnParams& bp, int prec, bool vir) {
    BpWork& w = work(b);
    const size_t rs = prec == P_FP64 ? 8 : 4;
    const int ldg = bp.nin_pad;
    // arrays sized for a larger list capacity serve a smaller one (the
    // kernels index them with the current capacity)
    const bool fits = w.a.G && w.cap >= b.cap && w.n == b.n && w.prec == prec && w.rev_elems >= b.nl_elems;
    if (!fits) {
        free_work(w);
        w.a.G = dnew<char>((size_t)b.n * ldg * rs);
        w.a.w = dnew<char>((size_t)b.n * ldg * rs);
        w.a.P = dnew<char>((size_t)b.n * b.cap * 3 * rs);
        w.a.own = dnew<char>((size_t)b.n * 3 * rs);
        w.a.rev = dnew<int>(b.nl_elems);
        w.cap = b.cap;
        w.n = b.n;
        w.prec = prec;
        w.rev_elems = b.nl_elems;
        w.rev_valid = false;
        b.alloc_gen++;
    }
    if (vir && !w.has_vir) {
        w.a.PV = dnew<char>((size_t)b.n * w.cap * 3 * rs);
        w.a.OV = dnew<char>((size_t)b.n * 3 * rs);
        w.has_vir = true;
        b.alloc_gen++;
    }
}

static void check_bpnn_box(const Batch& b) {
    for (int s = 0; s < b.nsys; ++s)
        for (int a = 0; a < 3; ++a)
            if (b.desc[s].per[a] && b.desc[s].L[a] < 2.0 * std::sqrt(b.desc[s].rns2))
                throw Error(E_INVAL, "BPNN needs periodic box lengths of at least twice the search radius");
    if (b.cap > 1024) throw Error(E_INVAL, "BPNN neighbour capacity above 1024");
}

// parts: 1 = descriptors and network (G, dE/dG, Ep), 2 = forces from the
// stored dE/dG, 4 (with 2) = the virial alone, added to V, force left as
// it is.  rev_fresh: the reverse-slot index in bufs is already built for
// the current list.
static void bpnn_run(Batch& b, const Pot& p, const BpBufs& W, bool rev_fresh, int what, int prec, int mode,
                     double dt, int step, const double4* pos, double4* pos_next, int parts) {
    const BpnnParams& bp = *p.bp;
    hipStream_t st = b.ctx->stream;
    // centres: the live records (the void tail has no list and no atoms)
    const int n = b.live;
    if (n == 0) return;
    const bool vir = (parts & 4) || (what & W_VIRIAL);
    const int vmode = (parts & 4) ? 2 : (vir ? 1 : 0);
    StepArgs S{dt, step, mode, what & (W_ENERGY | W_VIRIAL), mode == 0 ? 1 : 0};
    if (!rev_fresh) {
        const long long pairs = (long long)n * b.cap;
        hipLaunchKernelGGL(k_bp_rev, dim3((unsigned)((pairs + 255) / 256)), dim3(256), 0, st, b.nl, b.nn, W.rev, b.tpa,
                           b.M, b.cap, n);
    }
    const int ldg = bp.nin_pad;
    AtomArrays AA = arrays(b);
    double* Ep = b.d(b.sc, D_EP);
    const int4 boff = make_int4(bp.boff[0], bp.boff[1], bp.boff[2], bp.boff[3]);
    const int4 nts = make_int4(bp.nt[0], bp.nt[1], bp.nt[2], bp.nt[3]);
    const int4 kts = make_int4(bp.kt[0], bp.kt[1], bp.kt[2], bp.kt[3]);
    auto run = [&](auto rtag) {
        using R = decltype(rtag);
        // one wavefront per centre, four per workgroup unless the staged
        // neighbours of four take more LDS than a workgroup may hold
        const size_t lmax = std::min<size_t>(b.ctx->lds_max, 160 * 1024);
        int waves = 4;
        while (waves > 1 && waves * stage_bytes(b.cap, bp.h.ne, sizeof(R), true) > lmax) waves >>= 1;
        const size_t lds_desc = waves * stage_bytes(b.cap, bp.h.ne, sizeof(R), false);
        const size_t lds = waves * stage_bytes(b.cap, bp.h.ne, sizeof(R), true);
        if (lds > lmax) throw Error(E_INVAL, "BPNN neighbour staging exceeds the LDS");
        const dim3 gw((n + waves - 1) / waves), bw(64 * waves);
        // register arrays sized for the model: radial terms, (lambda, zeta)
        // groups and G4 eta values
        // 3, 4: the ladder kernels; 0-2: groups in order of appearance
        const int shape = bp.h.lz == 3 ? 3
                          : bp.h.lz == 5 ? 4
                          : (bp.h.nr <= 8 && bp.h.ng <= 4 && bp.h.ne <= 2) ? 0
                          : (bp.h.nr <= 8 ? 1 : 2);
        // K2: the two-species variants
        const bool two = bp.h.nspc == 2;
        auto pick_desc = [&](auto k2) {
            constexpr bool T2 = decltype(k2)::value;
            return shape == 3 ? k_bp_desc<R, 8, 6, 2, 3, T2>
                   : shape == 4 ? k_bp_desc<R, 8, 10, 7, 5, T2>
                   : shape == 0 ? k_bp_desc<R, 8, 4, 2, 0, T2>
                   : shape == 1 ? k_bp_desc<R, 8, kBpMaxG, kBpMaxEta, 0, T2>
                                : k_bp_desc<R, kBpMaxR, kBpMaxG, kBpMaxEta, 0, T2>;
        };
        auto kdesc = two ? pick_desc(std::true_type{}) : pick_desc(std::false_type{});
        // FP32, one species, descriptors through forces: the fused pass
        if constexpr (sizeof(R) == 4)
            if (!two && (parts & 3) == 3 && !(parts & 4)) {
                // ladder rungs with G4 terms: zeta in {1, 4} (rungs 0 and 2 of
                // three) and {1, 2, 4, 16} (all but rung 3 of five) have
                // kernels that skip the empty rungs
                int used = 0;
                for (int g = 0; g < kBpGridG; ++g)
                    for (int e = 0; e < kBpMaxEta; ++e)
                        if (bp.h.t_q[g * kBpMaxEta + e] >= 0) used |= 1 << (g >> 1);
                const bool z3 = (used & ~0b101) == 0, z5 = (used & ~0b10111) == 0;
                auto fused = [&](auto vtag) {
                    constexpr bool V = decltype(vtag)::value;
                    auto kf = shape == 3   ? (z3 ? k_bp_fused<8, 6, 2, V, 3, 0b101> : k_bp_fused<8, 6, 2, V, 3>)
                              : shape == 4 ? (z5 ? k_bp_fused<8, 10, 7, V, 5, 0b10111> : k_bp_fused<8, 10, 7, V, 5>)
                              : shape == 0 ? k_bp_fused<8, 4, 2, V, 0>
                              : shape == 1 ? k_bp_fused<8, kBpMaxG, kBpMaxEta, V, 0>
                                           : k_bp_fused<kBpMaxR, kBpMaxG, kBpMaxEta, V, 0>;
                    launch_k(b, K_BP_FUSED, kf, gw, bw, (unsigned)lds, pos, b.nl, b.nn, b.tpa, b.M, b.cap, b.d_desc,
                             bp.d, (const float*)bp.w64f, (float*)W.G, (float*)W.w, ldg, Ep, (float*)W.P,
                             (float*)W.own, (float*)W.PV, (float*)W.OV, n, S, b.misc);
                };
                if (vir) fused(std::true_type{});
                else fused(std::false_type{});
                parts &= ~3;
                parts |= 8;  // gather still to do
            }
        if (parts & 1) {
            launch_k(b, K_BP_DESC, kdesc, gw, bw, (unsigned)lds_desc, pos, b.nl, b.nn, b.tpa, b.M, b.cap, b.d_desc,
                     bp.d, (R*)W.G, ldg, n, S, b.misc);
            // one network per species, each over the atoms of its kind
            for (int sp = 0; sp < bp.h.nspc; ++sp) {
                const int only = bp.h.nspc > 1 ? sp : -1;
                if (sizeof(R) == 8) {
                    const bool narrow = std::max(bp.h.nin, std::max(bp.h.h1, bp.h.h2)) <= 32;
                    const int L = narrow ? 32 : 64;
                    const dim3 grid((unsigned)(((long long)n * L + 255) / 256));
                    launch_k(b, K_BP_MLP, narrow ? k_bp_mlp64<32> : k_bp_mlp64<64>, grid, dim3(256), 0,
                             (const double*)W.G, ldg, bp.w64 + sp * bp.w64_stride, bp.d, pos, (double*)W.w, Ep, n,
                             only, S, b.misc);
                } else
                    launch_k(b, K_BP_MLP, k_bp_mlp_mfma, dim3((n + 15) / 16), dim3(64), 0, (const float*)W.G, ldg,
                             bp.bsplit + sp * bp.b_stride, boff, nts, kts, bp.w32 + sp * bp.w32_stride, bp.d, pos,
                             (float*)W.w, Ep, n, only, S, b.misc);
            }
        }
        if (!(parts & 2) && !(parts & 8)) return;
        auto grad = [&](auto vtag) {
            constexpr bool V = decltype(vtag)::value;
            auto pick_grad = [&](auto k2) {
                constexpr bool T2 = decltype(k2)::value;
                return shape == 3 ? k_bp_grad<R, 8, 6, 2, V, 3, T2>
                       : shape == 4 ? k_bp_grad<R, 8, 10, 7, V, 5, T2>
                       : shape == 0 ? k_bp_grad<R, 8, 4, 2, V, 0, T2>
                       : shape == 1 ? k_bp_grad<R, 8, kBpMaxG, kBpMaxEta, V, 0, T2>
                                    : k_bp_grad<R, kBpMaxR, kBpMaxG, kBpMaxEta, V, 0, T2>;
            };
            auto kgrad = two ? pick_grad(std::true_type{}) : pick_grad(std::false_type{});
            launch_k(b, K_BP_FORCE, kgrad, gw, bw, (unsigned)lds, pos, b.nl, b.nn, b.tpa, b.M, b.cap, b.d_desc, bp.d,
                     (const R*)W.w, ldg, (R*)W.P, (R*)W.own, (R*)W.PV, (R*)W.OV, n, S, b.misc);
        };
        if (parts & 2) {
            if (vir) grad(std::true_type{});
            else grad(std::false_type{});
        }
        auto gather = [&](auto tpa_tag, auto vtag) {
            constexpr int T = decltype(tpa_tag)::value;
            constexpr bool V = decltype(vtag)::value;
            const long long waves = (n + kWave / T - 1) / (kWave / T);
            launch_k(b, K_BP_GATHER, k_bp_gather<R, T, V>, dim3((unsigned)((waves * kWave + 255) / 256)), dim3(256), 0,
                     pos, pos_next, b.nl, b.nn, W.rev, b.M, b.cap, b.d_desc, (const R*)W.P, (const R*)W.own,
                     (const R*)W.PV, (const R*)W.OV, vmode, AA, S, b.rsave, b.misc);
        };
        auto by_lanes = [&](auto vtag) {
            switch (b.tpa) {
                case 1: gather(std::integral_constant<int, 1>{}, vtag); break;
                case 2: gather(std::integral_constant<int, 2>{}, vtag); break;
                case 4: gather(std::integral_constant<int, 4>{}, vtag); break;
                case 8: gather(std::integral_constant<int, 8>{}, vtag); break;
                default: gather(std::integral_constant<int, 16>{}, vtag); break;
            }
        };
        if (vir) by_lanes(std::true_type{});
        else by_lanes(std::false_type{});
    };
    if (prec == P_FP64) run(double(0));
    else run(float(0));
}

// The work arrays carved one after the other from base (256-byte aligned
// each); returns the bytes used.  base may be null to count only.
static size_t carve_fixed(const Batch& b, const BpnnParams& bp, int prec, bool vir, char* base, BpBufs* out) {
    const size_t rs = prec == P_FP64 ? 8 : 4;
    size_t off = 0;
    auto take = [&](size_t bytes) -> void* {
        void* q = base ? base + off : nullptr;
        off += (bytes + 255) & ~(size_t)255;
        return q;
    };
    BpBufs W;
    W.G = take((size_t)b.n * bp.nin_pad * rs);
    W.w = take((size_t)b.n * bp.nin_pad * rs);
    W.P = take((size_t)b.n * b.cap * 3 * rs);
    W.own = take((size_t)b.n * 3 * rs);
    if (vir) {
        W.PV = take((size_t)b.n * b.cap * 3 * rs);
        W.OV = take((size_t)b.n * 3 * rs);
    }
    W.rev = static_cast<int*>(take(b.nl_elems * sizeof(int)));
    if (out) *out = W;
    return off;
}

size_t bpnn_fixed_bytes(const Batch& b, const Pot& p, int what, int prec) {
    return carve_fixed(b, *p.bp, prec, (what & W_VIRIAL) != 0, nullptr, nullptr);
}

void bpnn_compute(Batch& b, const Pot& p, int what, int prec, int mode, double dt, int step, const double4* pos,
                  double4* pos_next, int parts) {
    check_bpnn_box(b);
    if (b.bp_fixed) {
        // allocation-free: the work area is shared by every batch on the
        // device, so the reverse-slot index is built again for this list
        BpBufs W;
        carve_fixed(b, *p.bp, prec, (parts & 4) || (what & W_VIRIAL), b.bp_fixed, &W);
        bpnn_run(b, p, W, false, what, prec, mode, dt, step, pos, pos_next, parts);
        return;
    }
    const bool vir = (parts & 4) || (what & W_VIRIAL);
    ensure_work(b, *p.bp, prec, vir);
    BpWork& w = work(b);
    const bool fresh = w.rev_valid && w.rev_builds == b.rebuilds;
    bpnn_run(b, p, w.a, fresh, what, prec, mode, dt, step, pos, pos_next, parts);
    w.rev_valid = true;
    w.rev_builds = b.rebuilds;
}

// G and dE/dG of the last evaluation, original atom order, as doubles
void bpnn_inspect(Batch& b, const Pot& p, double* G, double* dEdG) {
    if (!b.bp_scratch) throw Error(E_STATE, "no BPNN evaluation to inspect");
    BpWork& w = work(b);
    const int ldg = p.bp->nin_pad, nin = p.bp->h.nin, n = b.n;
    const size_t rs = w.prec == P_FP64 ? 8 : 4;
    std::vector<char> g((size_t)n * ldg * rs), e((size_t)n * ldg * rs);
    std::vector<int> where(n);
    hipStream_t st = b.ctx->stream;
    HIPCHECK(hipMemcpyAsync(g.data(), w.a.G, g.size(), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(e.data(), w.a.w, e.size(), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(where.data(), b.where, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    for (int o = 0; o < n; ++o) {
        const int q = where[o];
        for (int k = 0; k < nin; ++k) {
            const size_t idx = (size_t)q * ldg + k;
            if (rs == 8) {
                if (G) G[(size_t)o * nin + k] = reinterpret_cast<double*>(g.data())[idx];
                if (dEdG) dEdG[(size_t)o * nin + k] = reinterpret_cast<double*>(e.data())[idx];
            } else {
                if (G) G[(size_t)o * nin + k] = reinterpret_cast<float*>(g.data())[idx];
                if (dEdG) dEdG[(size_t)o * nin + k] = reinterpret_cast<float*>(e.data())[idx];
            }
        }
    }
}

int bpnn_inputs(const Pot& p) { return p.bp ? p.bp->h.nin : 0; }

}  // namespace mdc
