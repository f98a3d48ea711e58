// Pair (LJsCut) and two-pass EAM force kernels with the velocity-Verlet
// update fused in, the standalone pre-drift kernel, the FP64 thermo
// reduction and the multi-step loop.
//
// Semantics restated from the reference:
//   LJsCut force/virial  solver/potential/LJs-cut.cxx:22-131
//   EAM density          solver/potential/EAM.cxx:499-560
//   EAM embedding        solver/potential/EAM.cxx:562-601 (once per atom here;
//                        the reference re-evaluates F'(rho_j) per pair)
//   EAM force/virial     solver/potential/EAM.cxx:306-497
//   Periodic ghost pass  solver/boundary/periodic.cxx:36-102 (pair_image)
//   velocity Verlet      solver/stepper/verlet.cxx:36-85
//   thermo               test/cxx/support/md_test_utils.h:214-239
//
// Precision: in the mixed mode the FP64 minimum-image difference is rounded
// to FP32 and the pair arithmetic and per-atom force accumulation run in
// FP32; energy and virial accumulate in FP64, and the EAM density sum,
// embedding, integration and displacement check stay FP64.
#include <climits>
#include <cmath>
#include "kern.h"
#include "rng.h"

namespace mdc {

// --------------------------------------------------------------------------
// Pair potentials: LJsCut, LJs, tabulated Pair
// --------------------------------------------------------------------------
struct LJK {
    double rc2;
    double a, s6, lj1, lj2, factor;                      // LJsCut
    double aVr, aVr2, rVr1, rVr2, aLJ, aLJ3, bLJ, bLJ2;  // LJs
    const double4* t64;                                  // Pair: r*phi records
    const float4* t32;
    unsigned nrec;
    int nr;
    double dr, dr1;
};

template <typename R>
__device__ inline typename Vec4T<R>::T load_rec(__amdgpu_buffer_rsrc_t b, int rec);
template <typename R>
__device__ inline const void* rtab_p(const double4* t64, const float4* t32) {
    return sizeof(R) == 8 ? (const void*)t64 : (const void*)t32;
}

// FN: 0 LJsCut (LJs-cut.cxx:22-131), 1 LJs (LJs.cxx:28-256), 2 tabulated
// Pair (pair.cxx:59-493).  Each gives V'(r)/r (the force on i along
// r_j - r_i) and V(r)/2 of one listed pair.
template <typename R, int TPA, bool EV, int FN>
__global__ void __launch_bounds__(256) k_lj(const double4* __restrict__ pos, double4* __restrict__ pos_next,
                                            const int* __restrict__ nl, const int* __restrict__ nn, int M, int cap,
                                            const SysDesc* __restrict__ desc, LJK P, AtomArrays A, StepArgs S,
                                            const double4* __restrict__ rsave, int* __restrict__ misc) {
    const AtomSlots as = atom_slots<TPA>(M);
    const int i = as.i;
    const bool valid = i < A.live;
    const int ab = abort_flag(S, misc);
    const double4 pi = valid ? pos[i] : make_double4(0, 0, 0, 0);
    const int nraw = valid ? nn[i] : 0;
    int j0[4];
    first_group<TPA, 4>(as, valid, M, nl, j0);
    if (ab <= S.step) return;
    const bool active = valid && !w_inactive(pi.w);
    const SysDesc& s = desc[valid ? w_sys(pi.w) : 0];
    const Frame fr = frame_of(s);
    const int cnt = active ? min(nraw, cap) : 0;
    const R c_pref = (R)(-0.5 * P.factor), c_lj1 = (R)P.lj1, c_lj2 = (R)P.lj2, c_s6 = (R)P.s6, c_e = (R)(0.5 * 4 * P.a);
    const R c_x2min = (R)1.2599210498948732, c_rVr1 = (R)P.rVr1, c_rVr2 = (R)P.rVr2, c_aVr = (R)P.aVr,
            c_aVr2 = (R)P.aVr2, c_aLJ = (R)P.aLJ, c_bLJ = (R)P.bLJ, c_aLJ3 = (R)P.aLJ3, c_bLJ2 = (R)P.bLJ2;
    const R c_dr = (R)P.dr, c_dr1 = (R)P.dr1;
    R fx = 0, fy = 0, fz = 0;
    double e = 0, vxx = 0, vyy = 0, vzz = 0, vxy = 0, vxz = 0, vyz = 0;
    const __amdgpu_buffer_rsrc_t posb = make_buf(pos, (unsigned)A.n * 32u);
    const __amdgpu_buffer_rsrc_t tabb =
        make_buf(FN == 2 ? rtab_p<R>(P.t64, P.t32) : (const void*)pos,
                 FN == 2 ? P.nrec * (unsigned)sizeof(typename Vec4T<R>::T) : 0u);
    for_neighbours<TPA, 4>(as, cnt, nl, posb, j0, [&](int ent, const double4& pj) {
        double dx, dy, dz, d2;
        if (!pair_delta(fr, pi, pj, ent, dx, dy, dz, d2)) return;
        if (d2 > P.rc2 || d2 < 1e-15) return;
        const R r2 = (R)d2, rx = (R)dx, ry = (R)dy, rz = (R)dz;
        R pref, eh = 0;
        if constexpr (FN == 0) {
            const R r2inv = rcp_(r2);
            const R r6inv = r2inv * r2inv * r2inv;
            pref = c_pref * (c_lj1 * r6inv - c_lj2) * r6inv * r2inv;
            if (EV) {
                const R t = c_s6 * r6inv;
                eh = c_e * t * (t - (R)1);
            }
        } else if constexpr (FN == 1) {
            const R XX = c_rVr1 * r2 - c_x2min;
            const R FV = c_rVr2 * rcp_(r2);
            const R Ur = FV * FV * FV;
            pref = c_aVr2 * (Ur * (Ur - (R)0.5) * FV + (c_aLJ + c_bLJ * XX) * XX);
            if (EV) eh = c_aVr * (Ur * (Ur - (R)1) - (c_aLJ3 + c_bLJ2 * XX) * XX * XX);
        } else {
            const R R1 = sqrt(r2), RN = R1 * c_dr1;
            const int N1 = (int)RN;
            if (N1 >= P.nr) return;  // beyond the table: V = V' = 0 (pair.cxx:370-376)
            R v, d;
            spline_vd<R>(load_rec<R>(tabb, N1), R1, RN, N1, c_dr, c_dr1, v, d);
            const R i1 = rcp_(R1);
            pref = (d * R1 - v) * rcp_(r2) * i1;
            if (EV) eh = (R)0.5 * v * i1;
        }
        fx += rx * pref;
        fy += ry * pref;
        fz += rz * pref;
        if (EV && (S.flags & W_ENERGY)) e += (double)eh;
        if (EV && (S.flags & W_VIRIAL)) {
            const double px = (double)(pref * rx), py = (double)(pref * ry), pz = (double)(pref * rz);
            vxx += px * (double)rx;
            vyy += py * (double)ry;
            vzz += pz * (double)rz;
            vxy += px * (double)ry;
            vxz += px * (double)rz;
            vyz += py * (double)rz;
        }
    });
    fx = wave_sum_group(fx, TPA);
    fy = wave_sum_group(fy, TPA);
    fz = wave_sum_group(fz, TPA);
    if (EV && (S.flags & W_ENERGY)) e = wave_sum_group(e, TPA);
    if (EV && (S.flags & W_VIRIAL)) {
        vxx = wave_sum_group(vxx, TPA);
        vyy = wave_sum_group(vyy, TPA);
        vzz = wave_sum_group(vzz, TPA);
        vxy = wave_sum_group(vxy, TPA);
        vxz = wave_sum_group(vxz, TPA);
        vyz = wave_sum_group(vyz, TPA);
    }
    if (!valid || as.sub != 0) return;
    if (!active) {
        if (S.mode == MODE_FINAL_DRIFT) pos_next[i] = pi;
        return;
    }
    const int n = A.n;
    if (EV && (S.flags & W_ENERGY)) A.d[(size_t)D_EP * n + i] = e;
    if (EV && (S.flags & W_VIRIAL)) {
        A.d[(size_t)D_VXX * n + i] = 0.5 * vxx;
        A.d[(size_t)D_VYY * n + i] = 0.5 * vyy;
        A.d[(size_t)D_VZZ * n + i] = 0.5 * vzz;
        A.d[(size_t)D_VXY * n + i] = 0.5 * vxy;
        A.d[(size_t)D_VXZ * n + i] = 0.5 * vxz;
        A.d[(size_t)D_VYZ * n + i] = 0.5 * vyz;
    }
    const FinIn fin = load_fin(i, A, S, rsave, true);
    finalize_atom<R>(i, pi, (double)fx, (double)fy, (double)fz, A, S, pos_next, fin, s, misc);
}

// --------------------------------------------------------------------------
// EAM: one element (EAM.cxx), or solver::Multi over one EAM object per pair
// of kinds (multi.cxx:11-151)
// --------------------------------------------------------------------------
struct EamK {
    int nk, nr, nrho, npair;
    unsigned nrec;  // interval records in r64 / r32
    double dr, dr1, drho, drho1, rc2;
    const double4* r64;
    const float4* r32;
    const float4* r32p;  // one element: rho and r*phi FP32 records of each interval side by side
    const double4* F64;
    const EamObj* objs;  // Multi: objects by pair_index(kind_i, kind_j)
    double* fpm;         // Multi: F' of every atom under the object of its kind and kind k, [nk][n]
    double4* posF;       // one element: (x, y, z, F'(nc)) of every atom, [n]
};

template <typename R>
__device__ inline typename Vec4T<R>::T load_rec(__amdgpu_buffer_rsrc_t b, int rec);
template <>
__device__ inline double4 load_rec<double>(__amdgpu_buffer_rsrc_t b, int rec) { return buf_d4(b, (unsigned)rec * 32u); }
template <>
__device__ inline float4 load_rec<float>(__amdgpu_buffer_rsrc_t b, int rec) { return buf_f4(b, (unsigned)rec * 16u); }
// the record, or zeros without a memory access for a pair that is not used
template <typename R>
__device__ inline typename Vec4T<R>::T load_rec_if(__amdgpu_buffer_rsrc_t b, int rec, bool ok) {
    return load_rec<R>(b, ok ? rec : (int)(kSkip / (unsigned)sizeof(typename Vec4T<R>::T)));
}
// The same inside neighbour loops that every lane runs alike: FP64 records
// (32 bytes) by the cooperative lane-pair load
template <typename R>
__device__ inline typename Vec4T<R>::T load_rec_lanes(__amdgpu_buffer_rsrc_t b, int rec, bool ok);
template <>
__device__ inline double4 load_rec_lanes<double>(__amdgpu_buffer_rsrc_t b, int rec, bool ok) {
    return buf_d4_pair(b, ok ? (unsigned)rec * 32u : kSkip);
}
template <>
__device__ inline float4 load_rec_lanes<float>(__amdgpu_buffer_rsrc_t b, int rec, bool ok) {
    return buf_f4(b, ok ? (unsigned)rec * 16u : kSkip);
}

template <typename R>
__device__ inline const typename Vec4T<R>::T* rtab(const EamK& P);
template <>
__device__ inline const double4* rtab<double>(const EamK& P) { return P.r64; }
template <>
__device__ inline const float4* rtab<float>(const EamK& P) { return P.r32; }

// Embedding F(rho) and F'(rho) in FP64 (EAM.cxx:562-589) on the interval
// records (y0, y0', y1, y1') of F at spacing h = drho: the cubic Hermite
// polynomial of the interval.  With x = rho - k h, t = x / h and the secant
// slope s = (y1 - y0) / h, and A = y0' + y1' - 2 s, B = 2 y0' + y1' - 3 s,
//   F = y0 + x (y0' + t (A t - B)),  F' = y0' + t (3 A t - 2 B).
// A density outside the table is flagged (the reference's branch for it is
// undefined) and yields NaN.
__device__ inline void embed_at(const double4* __restrict__ F64, int off, int nrho, double drho, double drho1,
                                double rho, double& F, double& dF, int* misc) {
    const double u = rho * drho1;
    const int k = (int)trunc(u);
    if (!(u >= 0.0 && k < nrho)) {
        atomicOr(&misc[M_ERR], EB_RHO_RANGE);
        F = dF = nan("");
        return;
    }
    const double4 rec = F64[off + k];  // y0, y0', y1, y1'
    const double t = u - (double)k;
    const double x = rho - (double)k * drho;
    const double sec = (rec.z - rec.x) * drho1;
    const double A = (rec.y - sec) + (rec.w - sec);
    const double B = A + (rec.y - sec);
    F = rec.x + x * (rec.y + t * (A * t - B));
    dF = rec.y + t * (3.0 * A * t - 2.0 * B);
}

// Pairs of one lane are processed kU at a time in predicated phases
// (distances, then every table and F' read, then arithmetic), so the kU
// dependent table reads are in flight together.  Two per group measured
// fastest at both precisions: four keep too many registers live (the FP64
// force kernel spills) and leave fewer waves to hide the gathers.
constexpr int kU = 2;

template <typename R>
struct PairGeo {
    double dx, dy, dz;  // FP64 difference (virial); rx, ry, rz: the same in R
    R rx, ry, rz;
    R R1, RN, R2;
    int N1;
    bool ok;
    bool ghost;  // the pair takes the Periodic boundary's ghost route (single-point evaluations)
};

template <typename R>
__device__ inline PairGeo<R> pair_geo(const Frame& s, const double4& pi, const double4& pj, int e, bool valid,
                                      double rc2, R dr1, int nr) {
    PairGeo<R> g;
    double d2;
    const bool img = valid && pair_delta(s, pi, pj, e, g.dx, g.dy, g.dz, d2);
    // periodic.cxx:36-60: a listed pair goes through a ghost when its raw
    // distance reaches max(rns) and its shortest image does not exceed it
    g.ghost = img && s.any_per && sq3_exact(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z) >= s.rns2;
    g.ok = img && !(d2 > rc2 || d2 < 1e-15);
    g.rx = (R)g.dx;
    g.ry = (R)g.dy;
    g.rz = (R)g.dz;
    g.R2 = g.ok ? (R)d2 : (R)1;
    g.R1 = sqrt(g.R2);
    g.RN = g.R1 * dr1;
    g.N1 = g.ok ? (int)g.RN : 0;
    if (g.N1 >= nr) {
        g.ok = false;
        g.N1 = 0;
    }
    return g;
}

__device__ inline int wave_max_group(int v, int tpa) {
    for (int o = tpa >> 1; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, kWave));
    return v;
}

// Multi: the objects are staged in LDS once per workgroup (before any
// early return, so every thread reaches the barrier).
template <bool MULTI>
__device__ inline void stage_objs(EamObj* s_obj, const EamK& P) {
    if (!MULTI) return;
    for (int q = threadIdx.x; q < P.npair; q += blockDim.x) s_obj[q] = P.objs[q];
    __syncthreads();
}

// Density pass (EAM.cxx:532-560), then F(rho) and F'(rho) per atom
// (EAM.cxx:562-601).  Multi: atom i sums the rho table of the object of
// (kind_i, kind_j) over each neighbour (multi.cxx:64-99) and keeps F' under
// every object of its kind; Em is F of the object of the last non-empty
// kind sub-list of the last pass over the atom's list (direct list, self
// included, then the ghost list when it is not empty), as the per-object
// virial() calls leave it (multi.cxx:128-151, EAM.cxx:306-354).
template <typename R, int TPA, bool MULTI>
__global__ void __launch_bounds__(256) k_eam_rho(const double4* __restrict__ pos, const int* __restrict__ nl,
                                                 const int* __restrict__ nn, int M, int cap,
                                                 const SysDesc* __restrict__ desc, EamK P, AtomArrays A, StepArgs S,
                                                 int* __restrict__ misc) {
    __shared__ EamObj s_obj[MULTI ? kMaxPairs : 1];
    stage_objs<MULTI>(s_obj, P);
    const AtomSlots as = atom_slots<TPA>(M);
    const int i = as.i;
    const bool valid = i < A.live;
    const int ab = abort_flag(S, misc);
    const double4 pi = valid ? pos[i] : make_double4(0, 0, 0, 0);
    const int nraw = valid ? nn[i] : 0;
    int j0[kU];
    first_group<TPA, kU>(as, valid, M, nl, j0);
    if (ab <= S.step) return;
    const bool active = valid && !w_inactive(pi.w);
    const SysDesc& s = desc[valid ? w_sys(pi.w) : 0];
    const Frame fr = frame_of(s);
    const int cnt = active ? min(nraw, cap) : 0;
    const auto* tab = rtab<R>(P);
    const R dr = (R)P.dr, dr1 = (R)P.dr1;
    const int ki = MULTI ? w_kind(pi.w) : 0;
    int kd = ki, kg = -1;  // Multi: largest kind in the direct list (self included) and among ghost pairs
    double rho = 0.0;
    const __amdgpu_buffer_rsrc_t posb = make_buf(pos, (unsigned)A.n * 32u);
    const __amdgpu_buffer_rsrc_t tabb = make_buf(tab, P.nrec * (unsigned)sizeof(*tab));
    int jn[kU];  // list entries of the next group, read one group ahead
#pragma unroll
    for (int u = 0; u < kU; ++u) jn[u] = j0[u];
    // every lane runs the wave's longest list (cooperative loads)
    const int nch = wave_chunks<TPA>(cnt);
    for (int m0 = 0; m0 < nch; m0 += kU) {
        int e[kU];
        bool okk[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int k = (m0 + u) * TPA + as.sub;
            okk[u] = k < cnt;
            e[u] = okk[u] ? jn[u] : 0;
        }
        double4 pj[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) pj[u] = buf_d4_pair(posb, okk[u] ? (unsigned)ent_j(e[u]) * 32u : kSkip);
        next_group<TPA, kU>(as, m0, M, nl, jn);
        PairGeo<R> g[kU];
        int ob[kU], toff[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (!MULTI) {
                ob[u] = toff[u] = 0;
                g[u] = pair_geo<R>(fr, pi, pj[u], e[u], okk[u], P.rc2, dr1, P.nr);
            } else {
                const int kj = w_kind(pj[u].w);
                ob[u] = pair_index(ki, kj);
                const EamObj& o = s_obj[ob[u]];
                toff[u] = ki > kj ? o.rho_hi_off : o.rho_off;  // the table of atom i's side
                g[u] = pair_geo<R>(fr, pi, pj[u], e[u], okk[u], o.rc2, (R)o.dr1, o.nr);
                // only an EAM object's virial sets Em (a Pair member's leaves it)
                if (okk[u] && !o.pair_only) kd = max(kd, kj);
                if (g[u].ghost && !o.pair_only) kg = max(kg, kj);
            }
        }
        typename Vec4T<R>::T rec[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) rec[u] = load_rec_lanes<R>(tabb, toff[u] + g[u].N1, g[u].ok);
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (!g[u].ok) continue;
            if (!MULTI) rho += (double)spline_v<R>(rec[u], g[u].R1, g[u].RN, g[u].N1, dr, dr1);
            else {
                const EamObj& o = s_obj[ob[u]];
                rho += (double)spline_v<R>(rec[u], g[u].R1, g[u].RN, g[u].N1, (R)o.dr, (R)o.dr1);
            }
        }
    }
    rho = wave_sum_group(rho, TPA);
    if (MULTI) {
        kd = wave_max_group(kd, TPA);
        kg = wave_max_group(kg, TPA);
    }
    if (!valid || as.sub != 0) return;
    const int n = A.n;
    // tag == 0: the reference skips the atom's prepare and charge, yet
    // neighbours still evaluate F'(nc) of its stored density
    // (single.cxx:44-50, EAM.cxx:467-474).  W_NC_STORED: a forces() or
    // virials() call on its own reads the density the atoms hold.
    const bool from_stored = !active || (S.flags & W_NC_STORED);
    const double rr = from_stored ? A.d[(size_t)D_NC * n + i] : rho;
    if (!MULTI) {
        double F, dF;
        embed_at(P.F64, 0, P.nrho, P.drho, P.drho1, rr, F, dF, misc);
        A.d[(size_t)D_FP * n + i] = dF;
        P.posF[i] = make_double4(pi.x, pi.y, pi.z, dF);
        if (active && S.store_aux && (S.flags & W_ENERGY)) A.d[(size_t)D_EM * n + i] = F;
    } else {
        const int klast = kg >= 0 ? kg : kd;
        // no EAM sub-list at all: Em stays as prepare left it (zero)
        if (klast < 0 && active && S.store_aux && (S.flags & W_ENERGY)) A.d[(size_t)D_EM * n + i] = 0.0;
        for (int k = 0; k < P.nk; ++k) {
            const EamObj& o = s_obj[pair_index(ki, k)];
            double F, dF;
            embed_at(P.F64, ki > k ? o.F_hi_off : o.F_off, o.nrho, o.drho, o.drho1, rr, F, dF, misc);
            P.fpm[(size_t)k * n + i] = dF;
            if (active && k == klast && S.store_aux && (S.flags & W_ENERGY)) A.d[(size_t)D_EM * n + i] = F;
        }
    }
    if (active && S.store_aux && !(S.flags & W_NC_STORED)) A.d[(size_t)D_NC * n + i] = rho;
}

// Force pass (EAM.cxx:356-497) and, when asked, pair energy and virial
// (EAM.cxx:306-354); the velocity-Verlet finish is fused.  Multi: each pair
// uses the object of (kind_i, kind_j): its rho'(r), its pair term and
// F'(nc_i) + F'(nc_j) under its own embedding table.  DIR (alloy EAM): the
// two sides of an object differ, so a pair of elements a != b takes
// F_a'(nc_i) rho_b'(r) + F_b'(nc_j) rho_a'(r), one more table record.
template <typename R, int TPA, bool MULTI, bool EV, bool DIR = false>
__global__ void __launch_bounds__(256) k_eam_force(const double4* __restrict__ pos, double4* __restrict__ pos_next,
                                                   const int* __restrict__ nl, const int* __restrict__ nn, int M,
                                                   int cap, const SysDesc* __restrict__ desc, EamK P, AtomArrays A,
                                                   StepArgs S, const double4* __restrict__ rsave,
                                                   int* __restrict__ misc) {
    __shared__ EamObj s_obj[MULTI ? kMaxPairs : 1];
    stage_objs<MULTI>(s_obj, P);
    const AtomSlots as = atom_slots<TPA>(M);
    const int i = as.i;
    const bool valid = i < A.live;
    const int n = A.n;
    const double* __restrict__ Fp = MULTI ? P.fpm : A.d + (size_t)D_FP * n;
    const int ab = abort_flag(S, misc);
    const double4 pi = valid ? pos[i] : make_double4(0, 0, 0, 0);
    const int nraw = valid ? nn[i] : 0;
    const int ki = MULTI ? w_kind(pi.w) : 0;
    double Fpk[MULTI ? kMaxMultiKinds : 1];  // F'(nc_i) under the object of (kind_i, k)
#pragma unroll
    for (int k = 0; k < (MULTI ? kMaxMultiKinds : 1); ++k)
        Fpk[k] = (valid && k < (MULTI ? P.nk : 1)) ? Fp[(size_t)k * n + i] : 0.0;
    int j0[kU];
    first_group<TPA, kU>(as, valid, M, nl, j0);
    if (ab <= S.step) return;
    const bool active = valid && !w_inactive(pi.w);
    const SysDesc& s = desc[valid ? w_sys(pi.w) : 0];
    const Frame fr = frame_of(s);
    const int cnt = active ? min(nraw, cap) : 0;
    const auto* tab = rtab<R>(P);
    const R dr = (R)P.dr, dr1 = (R)P.dr1;
    R fx = 0, fy = 0, fz = 0;
    double en = 0, vxx = 0, vyy = 0, vzz = 0, vxy = 0, vxz = 0, vyz = 0;
    // one element: neighbour position and F' come in one record (posF)
    const __amdgpu_buffer_rsrc_t posb = make_buf(MULTI ? (const void*)pos : (const void*)P.posF, (unsigned)n * 32u);
    const __amdgpu_buffer_rsrc_t fpb = make_buf(Fp, (unsigned)(MULTI ? P.nk : 1) * (unsigned)n * 8u);
    const __amdgpu_buffer_rsrc_t tabb = make_buf(tab, P.nrec * (unsigned)sizeof(*tab));
    // one element, FP32: each interval's rho and r*phi records as one 32-byte record
    const __amdgpu_buffer_rsrc_t tabpb = make_buf(P.r32p, (!MULTI && P.r32p) ? (unsigned)P.nr * 32u : 0u);
    const unsigned fp_base = (unsigned)ki * (unsigned)n;  // F' of neighbours under the objects of (k_j, kind_i)
    int jn[kU];  // list entries of the next group, read one group ahead
#pragma unroll
    for (int u = 0; u < kU; ++u) jn[u] = j0[u];
    // every lane runs the wave's longest list (cooperative loads)
    const int nch = wave_chunks<TPA>(cnt);
    for (int m0 = 0; m0 < nch; m0 += kU) {
        int e[kU];
        bool okk[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int k = (m0 + u) * TPA + as.sub;
            okk[u] = k < cnt;
            e[u] = okk[u] ? jn[u] : 0;
        }
        double4 pj[kU];
        R Fpj[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            pj[u] = buf_d4_pair(posb, okk[u] ? (unsigned)ent_j(e[u]) * 32u : kSkip);
            Fpj[u] = MULTI ? (R)buf_d1(fpb, okk[u] ? (fp_base + (unsigned)ent_j(e[u])) * 8u : kSkip) : (R)pj[u].w;
        }
        next_group<TPA, kU>(as, m0, M, nl, jn);
        PairGeo<R> g[kU];
        int ob[kU], kj[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (!MULTI) {
                ob[u] = kj[u] = 0;
                g[u] = pair_geo<R>(fr, pi, pj[u], e[u], okk[u], P.rc2, dr1, P.nr);
            } else {
                kj[u] = w_kind(pj[u].w);
                ob[u] = pair_index(ki, kj[u]);
                const EamObj& o = s_obj[ob[u]];
                g[u] = pair_geo<R>(fr, pi, pj[u], e[u], okk[u], o.rc2, (R)o.dr1, o.nr);
            }
        }
        typename Vec4T<R>::T rr[kU], rp[kU], rj[DIR ? kU : 1];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (!MULTI) {
                if constexpr (sizeof(R) == 4) {
                    buf_f4x2_pair(tabpb, g[u].ok ? (unsigned)g[u].N1 * 32u : kSkip, rr[u], rp[u]);
                } else {
                    rr[u] = load_rec_lanes<R>(tabb, g[u].N1, g[u].ok);
                    rp[u] = load_rec_lanes<R>(tabb, P.nr + g[u].N1, g[u].ok);
                }
            } else {
                const EamObj& o = s_obj[ob[u]];
                const bool hi = ki > kj[u];  // atom i is the higher kind of the pair
                rr[u] = load_rec_lanes<R>(tabb, (hi ? o.rho_hi_off : o.rho_off) + g[u].N1, g[u].ok);
                rp[u] = load_rec_lanes<R>(tabb, o.phi_off + g[u].N1, g[u].ok);
                if (DIR)
                    rj[u] = load_rec_lanes<R>(tabb, (hi ? o.rho_off : o.rho_hi_off) + g[u].N1, g[u].ok && kj[u] != ki);
            }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (!g[u].ok) continue;
            const R R1 = g[u].R1, RN = g[u].RN, R2 = g[u].R2;
            const int N1 = g[u].N1;
            R drp = dr, dr1p = dr1;
            double Fpi = Fpk[0];
            if (MULTI) {
                const EamObj& o = s_obj[ob[u]];
                drp = (R)o.dr;
                dr1p = (R)o.dr1;
#pragma unroll
                for (int k = 1; k < kMaxMultiKinds; ++k)
                    if (kj[u] == k) Fpi = Fpk[k];
            }
            R nd, rv, phv, phd;
            spline_vd<R>(rp[u], R1, RN, N1, drp, dr1p, phv, phd);
            spline_vd<R>(rr[u], R1, RN, N1, drp, dr1p, rv, nd);
            // FF = (AdF rho'(r) + phi'(r)) / r with phi' = ((r phi)' r - r phi) / r^2
            const R i1 = rcp_(R1), i2 = rcp_(R2);
            R FF;
            if (DIR && kj[u] != ki) {
                R rvj, ndj;
                spline_vd<R>(rj[u], R1, RN, N1, drp, dr1p, rvj, ndj);
                FF = ((R)(active ? Fpi : 0.0) * nd + Fpj[u] * ndj + (phd * R1 - phv) * i2) * i1;
            } else {
                const R AdF = (R)(active ? Fpi : 0.0) + Fpj[u];
                FF = (AdF * nd + (phd * R1 - phv) * i2) * i1;
            }
            if (EV && (S.flags & W_ENERGY)) en += 0.5 * (double)(phv * i1);
            const R rx = g[u].rx, ry = g[u].ry, rz = g[u].rz;
            fx += FF * rx;
            fy += FF * ry;
            fz += FF * rz;
            if (EV && (S.flags & W_VIRIAL)) {
                const double h = 0.5 * (double)FF;
                const double ax = h * g[u].dx, ay = h * g[u].dy, az = h * g[u].dz;
                vxx += ax * g[u].dx;
                vyy += ay * g[u].dy;
                vzz += az * g[u].dz;
                vxy += ax * g[u].dy;
                vxz += ax * g[u].dz;
                vyz += ay * g[u].dz;
            }
        }
    }
    fx = wave_sum_group(fx, TPA);
    fy = wave_sum_group(fy, TPA);
    fz = wave_sum_group(fz, TPA);
    if (EV && (S.flags & W_ENERGY)) en = wave_sum_group(en, TPA);
    if (EV && (S.flags & W_VIRIAL)) {
        vxx = wave_sum_group(vxx, TPA);
        vyy = wave_sum_group(vyy, TPA);
        vzz = wave_sum_group(vzz, TPA);
        vxy = wave_sum_group(vxy, TPA);
        vxz = wave_sum_group(vxz, TPA);
        vyz = wave_sum_group(vyz, TPA);
    }
    if (!valid || as.sub != 0) return;
    if (!active) {
        if (S.mode == MODE_FINAL_DRIFT) pos_next[i] = pi;
        return;
    }
    if (EV && (S.flags & W_ENERGY)) A.d[(size_t)D_EP * n + i] = en;
    if (EV && (S.flags & W_VIRIAL)) {
        A.d[(size_t)D_VXX * n + i] = vxx;
        A.d[(size_t)D_VYY * n + i] = vyy;
        A.d[(size_t)D_VZZ * n + i] = vzz;
        A.d[(size_t)D_VXY * n + i] = vxy;
        A.d[(size_t)D_VXZ * n + i] = vxz;
        A.d[(size_t)D_VYZ * n + i] = vyz;
    }
    const FinIn fin = load_fin(i, A, S, rsave, true);
    finalize_atom<R>(i, pi, (double)fx, (double)fy, (double)fz, A, S, pos_next, fin, s, misc);
}

// --------------------------------------------------------------------------
// Fixed-point step path: single-element EAM in mixed precision, every box
// with a unique minimum image between rebuilds (shift_ok).  Neighbours are
// gathered as one 16-byte fixed-point record (SysDesc qs) instead of a
// 32-byte FP64 position, and a pair's difference is the wrapped integer
// difference times qf: no image shift, no FP64 arithmetic per pair.  The
// formulas are those of k_eam_rho / k_eam_force (EAM.cxx:499-560 and
// 356-497) on the same FP32 interval records.
// --------------------------------------------------------------------------

// Slots of one lane per group of the fixed-point kernels (four measured no
// faster at 32 000 to a million atoms).
constexpr int kUq = 2;

// Pair geometry of the fixed-point path: difference, distance and table
// interval; ok is false for a pair beyond the cutoff or the table.
struct QGeo {
    float dx, dy, dz, r2, R1, RN;
    int N1;
    bool ok;
};
__device__ inline QGeo q_geo(const uint4& uj, const uint4& ui, bool listed, float qx, float qy, float qz, float rc2,
                             float dr1, int nr) {
#pragma clang fp contract(off)
    QGeo g;
    g.dx = fixq_d(uj.x, ui.x, qx);
    g.dy = fixq_d(uj.y, ui.y, qy);
    g.dz = fixq_d(uj.z, ui.z, qz);
    g.r2 = fmaf(g.dz, g.dz, fmaf(g.dy, g.dy, g.dx * g.dx));
    g.ok = listed && !(g.r2 > rc2 || g.r2 < 1e-15f);
    g.R1 = __builtin_amdgcn_sqrtf(g.r2);
    g.RN = g.R1 * dr1;
    g.N1 = (int)g.RN;
    g.ok = g.ok && g.N1 < nr;
    return g;
}

// Threads of the persistent LDS-table workgroups (one per CU).
constexpr int kQlThreads = 1024;

// Zeros the compiler cannot see through: the value of a masked-off table
// read.  With literal zeros it would fold the arithmetic on the record into
// the branch that reads it, and wait there for the read.
__device__ inline float4 opaque_zero4() {
    float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    asm volatile("" : "+v"(z.x), "+v"(z.y), "+v"(z.z), "+v"(z.w));
    return z;
}

// Fixed-point records of pos [0, live): after a sort, before a timing.
__global__ void __launch_bounds__(256) k_fixq(const double4* __restrict__ pos, uint4* __restrict__ uq,
                                              const SysDesc* __restrict__ desc, int live) {
    const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= live) return;
    const double4 p = pos[i];
    uq[i] = fixq4(p.x, p.y, p.z, p.w, desc[w_sys(p.w)]);
}

// The wavefront gw of a list packed with TPA lanes per atom (atom_slots for
// kernels that pick their wavefronts themselves).
template <int TPA>
__device__ inline AtomSlots atom_slots_w(long long gw, int M) {
    const int lane = threadIdx.x & (kWave - 1);
    AtomSlots a;
    a.i = (int)(gw * (kWave / TPA) + lane / TPA);
    a.sub = lane % TPA;
    a.base = gw * (long long)M * kWave + lane;
    return a;
}

// Table records of the fixed-point passes: from a copy of records [lo, nr)
// staged in LDS (LDS variants) or from global memory.  RPN: records per
// interval (1: rho; 2: rho and r*phi side by side).
template <bool LDS, int RPN>
__device__ inline void q_tab(__amdgpu_buffer_rsrc_t tb, const float4* s_tab, int lo, int N1, float4* rec) {
    if (LDS && N1 >= lo) {
#pragma unroll
        for (int k = 0; k < RPN; ++k) rec[k] = s_tab[RPN * (N1 - lo) + k];
    } else {
#pragma unroll
        for (int k = 0; k < RPN; ++k) rec[k] = buf_f4(tb, (unsigned)(RPN * N1 + k) * 16u);
    }
}

// The workgroup's least squared pair distance into misc[M_RMIN] when it is
// below the value the workgroup read (was) at its start: one atomic per
// workgroup, and none once the least distance is known (atomics of every
// wavefront on the one address serialised the pass).  Positive FP32 values
// order as their bit patterns do.
__device__ inline void q_note_rmin(float rmin, int was, int* __restrict__ misc) {
    __shared__ float s_min[kQlThreads / kWave];
    for (int o = kWave / 2; o > 0; o >>= 1) rmin = fminf(rmin, __shfl_xor(rmin, o, kWave));
    if (__lane_id() == 0) s_min[threadIdx.x / kWave] = rmin;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x / kWave); ++w) rmin = fminf(rmin, s_min[w]);
        if (__float_as_int(rmin) < was) atomicMin(&misc[M_RMIN], __float_as_int(rmin));
    }
}

// Neighbour records of the group of chunks [m0, m0 + kUq): entries jn,
// gathers of slots past the atom's count masked off (their records read as
// the atom's own).
template <int TPA>
__device__ inline void q_records(const AtomSlots& as, int m0, int cnt, __amdgpu_buffer_rsrc_t ub, const uint4& ui,
                                 const int* jn, bool* ok, uint4* uj) {
#pragma unroll
    for (int u = 0; u < kUq; ++u) {
        ok[u] = (m0 + u) * TPA + as.sub < cnt;
        uj[u] = ui;
        if (ok[u]) uj[u] = buf_u4(ub, (unsigned)ent_j(jn[u]) * 16u);
    }
}

// Density pass of one wavefront: rho_i from the rho records, then F(rho),
// F'(rho) in FP64 as k_eam_rho; writes the atom's record with F' in w for
// the force pass.
template <int TPA, bool ONE, bool LDS, bool PIPE>
__device__ inline void rho_q_wave(const AtomSlots& as, const uint4* __restrict__ uq, uint4* __restrict__ uqF,
                                  const int* __restrict__ nl, const int* __restrict__ nn, int M, int cap,
                                  const SysDesc* __restrict__ desc, const EamK& P, const AtomArrays& A,
                                  const StepArgs& S, int* __restrict__ misc, const float4* s_tab, int lo, int ab,
                                  float& rmin) {
    // every fused multiply-add written out: the LDS and global variants then
    // round alike, whatever the compiler would contract in each
#pragma clang fp contract(off)
    const int i = as.i;
    const bool valid = i < A.live;
    const uint4 ui = valid ? uq[i] : make_uint4(0, 0, 0, 0);
    const int nraw = valid ? nn[i] : 0;
    int j0[kUq];
    first_group<TPA, kUq>(as, valid, M, nl, j0);
    // the step's abort flag (read by the caller) is tested after the
    // prologue reads are issued, so that their latencies overlap
    if (ab <= S.step) return;
    const bool active = valid && !((ui.w >> 8) & 1u);
    const SysDesc& s = desc[ONE || !valid ? 0 : (int)(ui.w >> 10)];
    const float qx = s.qf[0], qy = s.qf[1], qz = s.qf[2];
    const int cnt = active ? min(nraw, cap) : 0;
    const float rc2 = (float)P.rc2, dr1 = (float)P.dr1;
    const int nr = P.nr;
    double rho = 0.0;
    const __amdgpu_buffer_rsrc_t ub = make_buf(uq, (unsigned)A.n * 16u);
    const __amdgpu_buffer_rsrc_t tabb = make_buf(P.r32, (unsigned)nr * 16u);
    const float4 z4 = opaque_zero4();
    int jn[kUq];
#pragma unroll
    for (int u = 0; u < kUq; ++u) jn[u] = j0[u];
    const int nch = wave_chunks<TPA>(cnt);
    // neighbour records of the first group, then of each next group one
    // group ahead (PIPE), so that their latency overlaps the table reads
    bool okk[kUq];
    uint4 uj[kUq];
    if (PIPE) {
        q_records<TPA>(as, 0, cnt, ub, ui, jn, okk, uj);
        next_group_clamped<kUq>(as, 0, M, nl, jn);
    }
    for (int m0 = 0; m0 < nch; m0 += kUq) {
        bool okn[kUq];
        uint4 ujn[kUq];
        // the gathers of unused slots are masked off
        q_records<TPA>(as, PIPE ? m0 + kUq : m0, cnt, ub, ui, jn, PIPE ? okn : okk, PIPE ? ujn : uj);
        next_group_clamped<kUq>(as, PIPE ? m0 + kUq : m0, M, nl, jn);
        QGeo g[kUq];
        float4 rec[kUq];
#pragma unroll
        for (int u = 0; u < kUq; ++u) {
            g[u] = q_geo(uj[u], ui, okk[u], qx, qy, qz, rc2, dr1, nr);
            rec[u] = z4;
            if (g[u].ok) q_tab<LDS, 1>(tabb, s_tab, lo, g[u].N1, &rec[u]);
            if (LDS) rmin = fminf(rmin, g[u].ok ? g[u].r2 : rmin);
        }
#pragma unroll
        for (int u = 0; u < kUq; ++u) {
            const float t = g[u].RN - (float)g[u].N1;
            const float4 c = rec[u];
            rho += (double)fmaf(t, fmaf(t, fmaf(t, c.w, c.z), c.y), c.x);
        }
        if (PIPE) {
#pragma unroll
            for (int u = 0; u < kUq; ++u) {
                okk[u] = okn[u];
                uj[u] = ujn[u];
            }
        }
    }
    rho = wave_sum_group(rho, TPA);
    if (!valid || as.sub != 0) return;
    const int n = A.n;
    // tag == 0: neighbours evaluate F'(nc) of its stored density
    const double rr = !active ? A.d[(size_t)D_NC * n + i] : rho;
    double F, dF;
    embed_at(P.F64, 0, P.nrho, P.drho, P.drho1, rr, F, dF, misc);
    A.d[(size_t)D_FP * n + i] = dF;
    uqF[i] = make_uint4(ui.x, ui.y, ui.z, __float_as_uint((float)dF));
    if (active && S.store_aux && (S.flags & W_ENERGY)) A.d[(size_t)D_EM * n + i] = F;
    if (active && S.store_aux) A.d[(size_t)D_NC * n + i] = rho;
}

// Force pass of one wavefront with the fused velocity-Verlet finish; also
// writes the drifted position's fixed-point record.
template <int TPA, bool ONE, bool LDS, bool PIPE>
__device__ inline void force_q_wave(const AtomSlots& as, const double4* __restrict__ pos,
                                    double4* __restrict__ pos_next, const uint4* __restrict__ uqF,
                                    uint4* __restrict__ uq_next, const int* __restrict__ nl,
                                    const int* __restrict__ nn, int M, int cap, const SysDesc* __restrict__ desc,
                                    const EamK& P, const AtomArrays& A, const StepArgs& S,
                                    const double4* __restrict__ rsave, int* __restrict__ misc, const float4* s_tab,
                                    int lo, int ab, float& rmin) {
#pragma clang fp contract(off)
    const int i = as.i;
    const bool valid = i < A.live;
    const int n = A.n;
    const double4 pi = valid ? pos[i] : make_double4(0, 0, 0, 0);
    const uint4 ui = valid ? uqF[i] : make_uint4(0, 0, 0, 0);
    const int nraw = valid ? nn[i] : 0;
    int j0[kUq];
    first_group<TPA, kUq>(as, valid, M, nl, j0);
    if (ab <= S.step) return;
    const bool active = valid && !w_inactive(pi.w);
    const SysDesc& s = desc[ONE || !valid ? 0 : w_sys(pi.w)];
    const float qx = s.qf[0], qy = s.qf[1], qz = s.qf[2];
    const int cnt = active ? min(nraw, cap) : 0;
    const float rc2 = (float)P.rc2, dr1 = (float)P.dr1;
    const int nr = P.nr;
    const float Fpi = active ? __uint_as_float(ui.w) : 0.f;
    // the finish's inputs are read now, so that their latency overlaps the
    // neighbour loop
    const FinIn fin = load_fin(i, A, S, rsave, valid);
    float fx = 0, fy = 0, fz = 0;
    const __amdgpu_buffer_rsrc_t ub = make_buf(uqF, (unsigned)n * 16u);
    // each interval's rho and r*phi records side by side (32 bytes)
    const __amdgpu_buffer_rsrc_t tabpb = make_buf(P.r32p, (unsigned)nr * 32u);
    const float4 z4 = opaque_zero4();
    int jn[kUq];
#pragma unroll
    for (int u = 0; u < kUq; ++u) jn[u] = j0[u];
    const int nch = wave_chunks<TPA>(cnt);
    bool okk[kUq];
    uint4 uj[kUq];
    if (PIPE) {
        q_records<TPA>(as, 0, cnt, ub, ui, jn, okk, uj);
        next_group_clamped<kUq>(as, 0, M, nl, jn);
    }
    for (int m0 = 0; m0 < nch; m0 += kUq) {
        bool okn[kUq];
        uint4 ujn[kUq];
        q_records<TPA>(as, PIPE ? m0 + kUq : m0, cnt, ub, ui, jn, PIPE ? okn : okk, PIPE ? ujn : uj);
        next_group_clamped<kUq>(as, PIPE ? m0 + kUq : m0, M, nl, jn);
        QGeo g[kUq];
        float4 rec[kUq][2];
#pragma unroll
        for (int u = 0; u < kUq; ++u) {
            g[u] = q_geo(uj[u], ui, okk[u], qx, qy, qz, rc2, dr1, nr);
            rec[u][0] = rec[u][1] = z4;
            if (g[u].ok) q_tab<LDS, 2>(tabpb, s_tab, lo, g[u].N1, rec[u]);
            if (LDS) rmin = fminf(rmin, g[u].ok ? g[u].r2 : rmin);
        }
#pragma unroll
        for (int u = 0; u < kUq; ++u) {
            const float t = g[u].RN - (float)g[u].N1, t3 = 3.f * t;
            const float4 a = rec[u][0], c = rec[u][1];
            // rho'(r), r*phi and (r*phi)'(r) of the interval cubic
            const float nd = fmaf(t, fmaf(t3, a.w, 2.f * a.z), a.y) * dr1;
            const float phv = fmaf(t, fmaf(t, fmaf(t, c.w, c.z), c.y), c.x);
            const float phd = fmaf(t, fmaf(t3, c.w, 2.f * c.z), c.y) * dr1;
            const float i1 = rcp_(g[u].R1), i2 = i1 * i1;
            const float AdF = Fpi + __uint_as_float(uj[u].w);
            // computed for every lane (the records of unused pairs read as
            // zeros), then selected: a conditional expression would become
            // a branch with the table reads sunk into it, one at a time
            float FF = fmaf(AdF, nd, fmaf(phd, g[u].R1, -phv) * i2) * i1;
            FF = g[u].ok ? FF : 0.f;
            fx = fmaf(FF, g[u].dx, fx);
            fy = fmaf(FF, g[u].dy, fy);
            fz = fmaf(FF, g[u].dz, fz);
        }
        if (PIPE) {
#pragma unroll
            for (int u = 0; u < kUq; ++u) {
                okk[u] = okn[u];
                uj[u] = ujn[u];
            }
        }
    }
    fx = wave_sum_group(fx, TPA);
    fy = wave_sum_group(fy, TPA);
    fz = wave_sum_group(fz, TPA);
    if (!valid || as.sub != 0) return;
    if (!active) {
        if (S.mode == MODE_FINAL_DRIFT) {
            pos_next[i] = pi;
            uq_next[i] = fixq4(pi.x, pi.y, pi.z, pi.w, s);
        }
        return;
    }
    finalize_atom<float>(i, pi, (double)fx, (double)fy, (double)fz, A, S, pos_next, fin, s, misc, uq_next);
}

// One wavefront per 64 / TPA atoms, tables read from global memory.
template <int TPA, bool ONE>
__global__ void __launch_bounds__(256) k_eam_rho_q(const uint4* __restrict__ uq, uint4* __restrict__ uqF,
                                                   const int* __restrict__ nl, const int* __restrict__ nn, int M,
                                                   int cap, const SysDesc* __restrict__ desc, EamK P, AtomArrays A,
                                                   StepArgs S, int* __restrict__ misc) {
    float rmin = 0.f;  // not tracked without staged tables
    rho_q_wave<TPA, ONE, false, true>(atom_slots<TPA>(M), uq, uqF, nl, nn, M, cap, desc, P, A, S, misc, nullptr, 0,
                                abort_flag(S, misc), rmin);
}

template <int TPA, bool ONE>
__global__ void __launch_bounds__(256) k_eam_force_q(const double4* __restrict__ pos, double4* __restrict__ pos_next,
                                                     const uint4* __restrict__ uqF, uint4* __restrict__ uq_next,
                                                     const int* __restrict__ nl, const int* __restrict__ nn, int M,
                                                     int cap, const SysDesc* __restrict__ desc, EamK P, AtomArrays A,
                                                     StepArgs S, const double4* __restrict__ rsave,
                                                     int* __restrict__ misc) {
    float rmin = 0.f;
    force_q_wave<TPA, ONE, false, true>(atom_slots<TPA>(M), pos, pos_next, uqF, uq_next, nl, nn, M, cap, desc, P, A, S,
                                  rsave, misc, nullptr, 0, abort_flag(S, misc), rmin);
}

// LDS variants: persistent workgroups of kQlThreads, one per CU, each of
// which stages the records [lo, nr) of the table once and then takes
// wavefronts of atoms in turn.  The neighbour gathers then leave the
// texture addresser, which bounds the passes with global table reads, to
// the list and position records.  Wavefronts are dealt per XCD: the
// workgroups of XCD x (blockIdx.x mod 8) take the x-th eighth of the
// cell-ordered wavefronts, whose neighbours stay in that XCD's L2.
__device__ inline void ql_range(long long nwaves, long long& first, long long& end, long long& step) {
    const int x = blockIdx.x % kXcd, k = blockIdx.x / kXcd, per = gridDim.x / kXcd;
    const int wpb = blockDim.x / kWave;
    first = nwaves * x / kXcd + (long long)k * wpb + threadIdx.x / kWave;
    end = nwaves * (x + 1) / kXcd;
    step = (long long)per * wpb;
}

// Stages the records [lo, nr) of the table and returns lo: the least
// interval a pair has met so far (misc[M_RMIN], less 2 %), but no lower than
// lo_cap, the least one the LDS can hold.  Pairs below lo read the table in
// global memory, so any lo gives the same results.
template <int RPN>
__device__ inline int ql_stage(float4* s_tab, const float4* __restrict__ tab, int lo_cap, int nr, float dr1,
                               const int* __restrict__ misc, int& was) {
    // one read for the workgroup: other workgroups lower the value meanwhile
    __shared__ int s_lo, s_was;
    if (threadIdx.x == 0) {
        const int rb = __atomic_load_n(&misc[M_RMIN], __ATOMIC_RELAXED);
        const float f = fminf(0.98f * __builtin_amdgcn_sqrtf(__int_as_float(rb)) * dr1, (float)nr);
        s_lo = max(lo_cap, (int)f);
        s_was = rb;
    }
    __syncthreads();
    const int lo = s_lo;
    was = s_was;
    const int nt = RPN * (nr - lo);
    for (int q = threadIdx.x; q < nt; q += blockDim.x) s_tab[q] = tab[RPN * lo + q];
    __syncthreads();
    return lo;
}

template <int TPA, bool ONE>
__global__ void __launch_bounds__(kQlThreads) k_eam_rho_ql(const uint4* __restrict__ uq, uint4* __restrict__ uqF,
                                                          const int* __restrict__ nl, const int* __restrict__ nn,
                                                          int M, int cap, const SysDesc* __restrict__ desc, EamK P,
                                                          AtomArrays A, StepArgs S, int* __restrict__ misc, int lo,
                                                          long long nwaves) {
    extern __shared__ float4 s_tab[];
    if (abort_flag(S, misc) <= S.step) return;
    int was;
    lo = ql_stage<1>(s_tab, P.r32, lo, P.nr, (float)P.dr1, misc, was);
    long long gw, end, step;
    ql_range(nwaves, gw, end, step);
    float rmin = __int_as_float(kRminNone);
    for (; gw < end; gw += step)
        rho_q_wave<TPA, ONE, true, false>(atom_slots_w<TPA>(gw, M), uq, uqF, nl, nn, M, cap, desc, P, A, S, misc, s_tab,
                                   lo, INT_MAX, rmin);
    q_note_rmin(rmin, was, misc);
}

template <int TPA, bool ONE>
__global__ void __launch_bounds__(kQlThreads) k_eam_force_ql(
    const double4* __restrict__ pos, double4* __restrict__ pos_next, const uint4* __restrict__ uqF,
    uint4* __restrict__ uq_next, const int* __restrict__ nl, const int* __restrict__ nn, int M, int cap,
    const SysDesc* __restrict__ desc, EamK P, AtomArrays A, StepArgs S, const double4* __restrict__ rsave,
    int* __restrict__ misc, int lo, long long nwaves) {
    extern __shared__ float4 s_tab[];
    if (abort_flag(S, misc) <= S.step) return;
    int was;
    lo = ql_stage<2>(s_tab, P.r32p, lo, P.nr, (float)P.dr1, misc, was);
    long long gw, end, step;
    ql_range(nwaves, gw, end, step);
    float rmin = __int_as_float(kRminNone);
    for (; gw < end; gw += step)
        force_q_wave<TPA, ONE, true, false>(atom_slots_w<TPA>(gw, M), pos, pos_next, uqF, uq_next, nl, nn, M, cap, desc,
                                     P, A, S, rsave, misc, s_tab, lo, INT_MAX, rmin);
    q_note_rmin(rmin, was, misc);
}

// Opening half kick + drift of the first step of a sequence from the stored
// force (verlet.cxx:71-78); the trigger marks step 0 as needing a rebuild.
// uq_next (fixed-point path, else null): the drifted position's
// fixed-point record.
__global__ void k_predrift(const double4* __restrict__ pos, double4* __restrict__ pos_next, AtomArrays A, int f_prec,
                           double dt, const SysDesc* __restrict__ desc, const double4* __restrict__ rsave,
                           int* __restrict__ misc, uint4* __restrict__ uq_next) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = A.n;
    if (i >= A.live) return;
    const double4 pi = pos[i];
    if (w_inactive(pi.w)) {
        pos_next[i] = pi;
        if (uq_next) uq_next[i] = fixq4(pi.x, pi.y, pi.z, pi.w, desc[w_sys(pi.w)]);
        return;
    }
    double F[3];
    if (f_prec == P_MIXED) {
        F[0] = A.f[(size_t)F_FX * n + i];
        F[1] = A.f[(size_t)F_FY * n + i];
        F[2] = A.f[(size_t)F_FZ * n + i];
    } else {
        F[0] = A.d[(size_t)D_FX * n + i];
        F[1] = A.d[(size_t)D_FY * n + i];
        F[2] = A.d[(size_t)D_FZ * n + i];
    }
    const double m = A.d[(size_t)D_M * n + i];
    double r[3] = {pi.x, pi.y, pi.z};
    for (int c = 0; c < 3; ++c) {
        double a = F[c] / m;
        double v = A.d[(size_t)(D_VX + c) * n + i] + a * dt * 0.5;
        r[c] += v * dt;
        A.d[(size_t)(D_VX + c) * n + i] = v;
    }
    pos_next[i] = make_double4(r[0], r[1], r[2], pi.w);
    const SysDesc& s = desc[w_sys(pi.w)];
    if (uq_next) uq_next[i] = fixq4(r[0], r[1], r[2], pi.w, s);
    const double4 q = rsave[i];
    double dx = r[0] - q.x, dy = r[1] - q.y, dz = r[2] - q.z;
    shortest3(s, dx, dy, dz);
    wave_atomic_min(sq3_exact(dx, dy, dz) > s.trig2, &misc[M_ABORT], 0);
}

// Per-system FP64 reduction in a fixed order (one workgroup per system):
// out[s] = [Ep, Em, KE, Etot, W, com_x, com_y, com_z].
// One workgroup per box over the box's atoms, [off[2 s], off[2 s] +
// off[2 s + 1]) of the caller's order, each at its sorted place where[o].
__global__ void __launch_bounds__(256) k_thermo(const double* __restrict__ d, int n, const int* __restrict__ off,
                                                const int* __restrict__ where, double* __restrict__ out) {
    const int sys = blockIdx.x;
    const int b = off[2 * sys], e = b + off[2 * sys + 1];
    __shared__ double red[8][256];
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // Ep, Em, W, px, py, pz, M
    for (int o = b + threadIdx.x; o < e; o += 256) {
        const int i = where[o];
        acc[0] += d[(size_t)D_EP * n + i];
        acc[1] += d[(size_t)D_EM * n + i];
        acc[2] += d[(size_t)D_VXX * n + i] + d[(size_t)D_VYY * n + i] + d[(size_t)D_VZZ * n + i];
        const double m = d[(size_t)D_M * n + i];
        acc[3] += m * d[(size_t)D_VX * n + i];
        acc[4] += m * d[(size_t)D_VY * n + i];
        acc[5] += m * d[(size_t)D_VZ * n + i];
        acc[6] += m;
    }
    for (int k = 0; k < 7; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int k = 0; k < 7; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + o];
        __syncthreads();
    }
    const double M = red[6][0];
    double com[3] = {0, 0, 0};
    if (M > 0.0)
        for (int c = 0; c < 3; ++c) com[c] = red[3 + c][0] / M;
    __syncthreads();
    double ke = 0.0;
    for (int o = b + threadIdx.x; o < e; o += 256) {
        const int i = where[o];
        const double m = d[(size_t)D_M * n + i];
        double rx = d[(size_t)D_VX * n + i] - com[0], ry = d[(size_t)D_VY * n + i] - com[1],
               rz = d[(size_t)D_VZ * n + i] - com[2];
        ke += 0.5 * m * (rx * rx + ry * ry + rz * rz);
    }
    red[7][threadIdx.x] = ke;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[7][threadIdx.x] += red[7][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* o = out + 8 * sys;
        o[0] = red[0][0];
        o[1] = red[1][0];
        o[2] = red[7][0];
        o[3] = red[0][0] + red[1][0] + red[7][0];
        o[4] = red[2][0];
        o[5] = com[0];
        o[6] = com[1];
        o[7] = com[2];
    }
}

// --------------------------------------------------------------------------
// host side
// --------------------------------------------------------------------------
static inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

hipEvent_t timer_event(Batch& b) {
    hipEvent_t e;
    if (!b.ev_free.empty()) {
        e = b.ev_free.back();
        b.ev_free.pop_back();
    } else {
        HIPCHECK(hipEventCreate(&e));
    }
    return e;
}

// Marker-event timing of a sequence of launches (the list rebuild).
hipEvent_t timer_begin(Batch& b) {
    if (!b.timing) return nullptr;
    hipEvent_t e = timer_event(b);
    HIPCHECK(hipEventRecord(e, b.ctx->stream));
    return e;
}

void timer_end(Batch& b, hipEvent_t a, int kind) {
    if (!b.timing || !a) return;
    hipEvent_t e = timer_event(b);
    HIPCHECK(hipEventRecord(e, b.ctx->stream));
    b.ev_pending.push_back({a, e, kind});
}

void timer_drain(Batch& b) {
    for (auto& t : b.ev_pending) {
        float ms = 0;
        HIPCHECK(hipEventSynchronize(t.b));
        HIPCHECK(hipEventElapsedTime(&ms, t.a, t.b));
        b.kt_ms[t.kind] += ms;
        b.kt_n[t.kind] += 1;
        b.ev_free.push_back(t.a);
        b.ev_free.push_back(t.b);
    }
    b.ev_pending.clear();
}

void set_trigger(Batch& b, double rcut, bool allow_shift) {
    bool changed = false;
    b.all_shift = b.nsys > 0;
    for (int s = 0; s < b.nsys; ++s) {
        SysDesc& d = b.desc[s];
        double skin = b.rns_min[s] - rcut;
        double t = skin > 0.0 ? 0.25 * skin * skin : -1.0;
        // within a step sequence every atom stays within skin/2 of its
        // position at the last build, so listed pairs stay within
        // rns_max + skin; their minimum image is unique while that is below
        // half of every periodic edge
        int sh = allow_shift && skin > 0.0 && !b.shifts_stale ? 1 : 0;
        for (int a = 0; a < 3; ++a)
            if (d.per[a] && !(0.5 * d.L[a] > b.rns_max[s] + skin)) sh = 0;
        if (allow_shift && b.manual_rebuild) {
            // the caller rebuilds: no trigger, and each pair's image is
            // chosen from the current positions at every step
            t = 1e300;
            sh = 0;
        }
        if (d.trig2 != t || d.shift_ok != sh) {
            d.trig2 = t;
            d.shift_ok = sh;
            changed = true;
        }
        if (!sh) b.all_shift = false;
        // fixed-point scales: listed pairs stay within rns_max + skin, below
        // half of every periodic edge (shift_ok) and below a quarter of the
        // power-of-two modulus of an open axis
        for (int a = 0; a < 3 && sh; ++a) {
            double mod = d.per[a] ? d.L[a] : std::exp2(std::ceil(std::log2(4.0 * (b.rns_max[s] + skin))));
            if (!(mod > 0.0 && mod < 1e300)) mod = 1.0;  // a box with no atoms
            const double qs = std::ldexp(1.0, 32) / mod;
            const float qf = (float)(mod * std::ldexp(1.0, -32));
            if (d.qs[a] != qs || d.qf[a] != qf) {
                d.qs[a] = qs;
                d.qf[a] = qf;
                changed = true;
            }
        }
    }
    if (changed)
        HIPCHECK(hipMemcpyAsync(b.d_desc, b.desc.data(), sizeof(SysDesc) * b.nsys, hipMemcpyHostToDevice,
                                b.ctx->stream));
}

static EamK eam_args(const Pot& p, Batch& b) {
    EamK k;
    k.nk = p.eam.nk;
    k.nr = p.eam.nr;
    k.nrho = p.eam.nrho;
    k.npair = p.eam.npair;
    k.nrec = (unsigned)p.eam.nrec;
    k.dr = p.eam.dr;
    k.dr1 = 1.0 / p.eam.dr;
    k.drho = p.eam.drho;
    k.drho1 = 1.0 / p.eam.drho;
    k.rc2 = p.eam.rc * p.eam.rc;
    k.r64 = p.eam.r64;
    k.r32p = p.eam.r32p;
    k.r32 = p.eam.r32;
    k.F64 = p.eam.F64;
    k.objs = p.eam.objs;
    k.fpm = nullptr;
    if (p.eam.nk > 1) {
        const size_t need = (size_t)p.eam.nk * (size_t)b.n;
        if (b.fpm_elems < need) {
            if (b.fpm) HIPCHECK(hipFree(b.fpm));
            b.fpm = nullptr;
            b.fpm_elems = 0;
            if (hipMalloc(&b.fpm, need * sizeof(double)) != hipSuccess)
                throw Error(E_NOMEM, "out of device memory for the Multi embedding derivatives");
            b.fpm_elems = need;
            b.alloc_gen++;
        }
        k.fpm = b.fpm;
    }
    k.posF = b.posF;
    return k;
}

static LJK lj_args(const Pot& p) {
    const LJParams& L = p.lj;
    LJK k{};
    k.rc2 = L.rc2;
    k.a = L.a;
    k.s6 = L.s6;
    k.lj1 = L.lj1;
    k.lj2 = L.lj2;
    k.factor = L.factor;
    k.aVr = L.aVr;
    k.aVr2 = L.aVr2;
    k.rVr1 = L.rVr1;
    k.rVr2 = L.rVr2;
    k.aLJ = L.aLJ;
    k.aLJ3 = L.aLJ3;
    k.bLJ = L.bLJ;
    k.bLJ2 = L.bLJ2;
    k.t64 = p.eam.r64;
    k.t32 = p.eam.r32;
    k.nrec = (unsigned)p.eam.nrec;
    k.nr = p.eam.nr;
    k.dr = p.eam.dr;
    k.dr1 = p.eam.dr > 0 ? 1.0 / p.eam.dr : 0.0;
    return k;
}

// which: 1 = density pass (EAM), 2 = force pass, 3 = both
template <typename R, int TPA>
static void launch_pair_t(Batch& b, const Pot& p, const StepArgs& S, const double4* pos, double4* pos_next,
                          int which) {
    hipStream_t st = b.ctx->stream;
    if (b.live == 0) return;
    const int A = kWave / TPA;
    const long long waves = (b.live + A - 1) / A;
    const int grid = nblk(waves * kWave, b.block);
    AtomArrays AA = arrays(b);
    if (p.type == PT_LJ) {
        if (!(which & 2)) return;
        // energy and virial code only in the variant that computes them: the
        // step kernels keep fewer registers live
        const bool ev = S.flags & (W_ENERGY | W_VIRIAL);
        auto pick = [&](auto fn) {
            constexpr int F = decltype(fn)::value;
            return ev ? k_lj<R, TPA, true, F> : k_lj<R, TPA, false, F>;
        };
        auto k = p.lj.fn == 1 ? pick(std::integral_constant<int, 1>{})
                              : p.lj.fn == 2 ? pick(std::integral_constant<int, 2>{})
                                             : pick(std::integral_constant<int, 0>{});
        launch_k(b, K_FORCE, k, dim3(grid), dim3(b.block), 0, pos, pos_next, b.nl, b.nn, b.M, b.cap, b.d_desc,
                 lj_args(p), AA, S, b.rsave, b.misc);
    } else {
        EamK E = eam_args(p, b);
        if (sizeof(R) == 4 && b.qfast) {
            // fixed-point step path: uq[c] holds the records of pos = pos[c]
            const int c = pos == b.pos[0] ? 0 : 1;
            // one box: its descriptor by uniform loads at the kernel's start
            const bool one = b.nsys == 1;
            uint4* uq_next = pos_next ? b.uq[c ^ 1] : (uint4*)nullptr;
            if (b.q_lds) {
                // persistent workgroups, one per CU, each staging the table
                // records [lo, nr) that the LDS holds
                // (less a little for the kernels' own LDS variables)
                const size_t lmax = std::min<size_t>(b.ctx->lds_max, 160 * 1024) - 256;
                const int nr = p.eam.nr;
                const int lo1 = std::max(0, nr - (int)(lmax / 16)), lo2 = std::max(0, nr - (int)(lmax / 32));
                const int wpb = kQlThreads / kWave;
                int G = (int)std::min<long long>(b.ctx->cu_count, (waves + wpb - 1) / wpb);
                G = std::max(kXcd, (G + kXcd - 1) / kXcd * kXcd);
                if (which & 1)
                    launch_k(b, K_DENSITY, one ? k_eam_rho_ql<TPA, true> : k_eam_rho_ql<TPA, false>, dim3(G),
                             dim3(kQlThreads), (unsigned)(16 * (nr - lo1)), (const uint4*)b.uq[c], b.uqF, b.nl, b.nn,
                             b.M, b.cap, b.d_desc, E, AA, S, b.misc, lo1, waves);
                if (which & 2)
                    launch_k(b, K_FORCE, one ? k_eam_force_ql<TPA, true> : k_eam_force_ql<TPA, false>, dim3(G),
                             dim3(kQlThreads), (unsigned)(32 * (nr - lo2)), pos, pos_next, (const uint4*)b.uqF,
                             uq_next, b.nl, b.nn, b.M, b.cap, b.d_desc, E, AA, S, b.rsave, b.misc, lo2, waves);
                return;
            }
            if (which & 1)
                launch_k(b, K_DENSITY, one ? k_eam_rho_q<TPA, true> : k_eam_rho_q<TPA, false>, dim3(grid),
                         dim3(b.block), 0, (const uint4*)b.uq[c], b.uqF, b.nl, b.nn, b.M, b.cap, b.d_desc, E, AA, S,
                         b.misc);
            if (which & 2)
                launch_k(b, K_FORCE, one ? k_eam_force_q<TPA, true> : k_eam_force_q<TPA, false>, dim3(grid),
                         dim3(b.block), 0, pos, pos_next, (const uint4*)b.uqF, uq_next, b.nl, b.nn, b.M, b.cap,
                         b.d_desc, E, AA, S, b.rsave, b.misc);
            return;
        }
        if (which & 1) {
            if (E.nk > 1)
                launch_k(b, K_DENSITY, k_eam_rho<R, TPA, true>, dim3(grid), dim3(b.block), 0, pos, b.nl, b.nn, b.M, b.cap,
                         b.d_desc, E, AA, S, b.misc);
            else
                launch_k(b, K_DENSITY, k_eam_rho<R, TPA, false>, dim3(grid), dim3(b.block), 0, pos, b.nl, b.nn, b.M,
                         b.cap, b.d_desc, E, AA, S, b.misc);
        }
        if (!(which & 2)) return;
        const bool ev = S.flags & (W_ENERGY | W_VIRIAL);
        auto k = p.eam.alloy ? (ev ? k_eam_force<R, TPA, true, true, true> : k_eam_force<R, TPA, true, false, true>)
                 : E.nk > 1  ? (ev ? k_eam_force<R, TPA, true, true> : k_eam_force<R, TPA, true, false>)
                             : (ev ? k_eam_force<R, TPA, false, true> : k_eam_force<R, TPA, false, false>);
        launch_k(b, K_FORCE, k, dim3(grid), dim3(b.block), 0, pos, pos_next, b.nl, b.nn, b.M, b.cap, b.d_desc, E, AA, S,
                 b.rsave, b.misc);
    }
}

template <typename R>
static void launch_pair_r(Batch& b, const Pot& p, const StepArgs& S, const double4* pos, double4* pos_next,
                          int which = 3) {
    switch (b.tpa) {
        case 1: launch_pair_t<R, 1>(b, p, S, pos, pos_next, which); break;
        case 2: launch_pair_t<R, 2>(b, p, S, pos, pos_next, which); break;
        case 4: launch_pair_t<R, 4>(b, p, S, pos, pos_next, which); break;
        case 8: launch_pair_t<R, 8>(b, p, S, pos, pos_next, which); break;
        case 16: launch_pair_t<R, 16>(b, p, S, pos, pos_next, which); break;
        default: throw Error(E_INVAL, "lanes per atom must be 1, 2, 4, 8 or 16");
    }
}

static void launch_step_kernels(Batch& b, const Pot& p, const StepArgs& S, int prec, const double4* pos,
                                double4* pos_next) {
    if (p.type == PT_BPNN) {
        bpnn_compute(b, p, S.flags, prec, S.mode, S.dt, S.step, pos, pos_next);
        return;
    }
    if (prec == P_MIXED) launch_pair_r<float>(b, p, S, pos, pos_next);
    else launch_pair_r<double>(b, p, S, pos, pos_next);
}

void check_errors(Batch& b, const int* counters) {
    int e = (counters ? counters : b.ctx->h_pin)[M_ERR];
    if (!e) return;
    HIPCHECK(hipMemsetAsync(b.misc + M_ERR, 0, sizeof(int), b.ctx->stream));
    if (e & EB_RHO_RANGE)
        throw Error(E_RANGE, "EAM density outside the embedding table (the reference's branch for it is undefined)");
    if (e & EB_NONFINITE) throw Error(E_RANGE, "non-finite position after integration");
    throw Error(E_RANGE, "device error flag set");
}

static void read_misc(Batch& b) {
    HIPCHECK(hipMemcpyAsync(b.ctx->h_pin, b.misc, sizeof(int) * M_COUNT, hipMemcpyDeviceToHost, b.ctx->stream));
    stream_wait(*b.ctx);
    timer_drain(b);
}

// --------------------------------------------------------------------------
// Langevin thermostat (thermostat/langevin.cxx:29-42, applied after the
// atom's force in pair.cxx:72-87 and ternary.cxx:42-55)
// --------------------------------------------------------------------------
// f += ((sigma * g) sqrt(m) - (v - U) beta m) * heat for every thermostat
// whose [xmin, xmax] holds the atom's x, with g three normal deviates of
// the counter-based stream at (atom, call, thermostat).
__device__ inline void langevin_add(const ThermArgs& T, uint64_t call, int atom, double x, const double v[3],
                                    double m, double f[3]) {
    const double sm = sqrt(m);
    for (int k = 0; k < T.count; ++k) {
        const LangevinP& L = T.t[k];
        if (x < L.xmin || x > L.xmax) continue;
        double g[3];
        normal3(T.seed, call, (uint32_t)atom, k, g);
        for (int c = 0; c < 3; ++c) {
            double t = (L.sigma * g[c]) * sm;
            t -= (v[c] - L.U[c]) * L.beta * m;
            t *= L.heat[c];
            f[c] += t;
        }
    }
}

// A force evaluation outside a step sequence: add the thermostat force to
// the stored force of precision `prec`, with the stored velocity.
__global__ void k_langevin_apply(const double4* __restrict__ pos, AtomArrays A, const int* __restrict__ orig,
                                 ThermArgs T, uint64_t call, int prec) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = A.n;
    if (i >= A.live) return;
    const double4 pi = pos[i];
    if (w_inactive(pi.w)) return;
    double f[3], v[3];
    for (int c = 0; c < 3; ++c) {
        f[c] = prec == P_MIXED ? (double)A.f[(size_t)(F_FX + c) * n + i] : A.d[(size_t)(D_FX + c) * n + i];
        v[c] = A.d[(size_t)(D_VX + c) * n + i];
    }
    langevin_add(T, call, orig[i], pi.x, v, A.d[(size_t)D_M * n + i], f);
    for (int c = 0; c < 3; ++c) {
        if (prec == P_MIXED) A.f[(size_t)(F_FX + c) * n + i] = (float)f[c];
        else A.d[(size_t)(D_FX + c) * n + i] = f[c];
    }
}

// Step t of a sequence with thermostats: the force pass stored the
// potential force (MODE_FORCE_STEP); add the thermostat force at the
// half-step velocity and the drifted position (verlet.cxx:71-85: the
// solver's forces run between the drift and the closing kick), then the
// closing kick and, unless last, the next opening kick and drift.
template <typename R>
__global__ void k_langevin_finish(const double4* __restrict__ pos, double4* __restrict__ pos_next, AtomArrays A,
                                  StepArgs S, const int* __restrict__ orig, ThermArgs T, uint64_t call,
                                  const SysDesc* __restrict__ desc, const double4* __restrict__ rsave,
                                  int* __restrict__ misc) {
    if (skip_step(S, misc)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = A.n;
    if (i >= A.live) return;
    const double4 pi = pos[i];
    if (w_inactive(pi.w)) {
        if (S.mode == MODE_FINAL_DRIFT) pos_next[i] = pi;
        return;
    }
    const FinIn fin = load_fin(i, A, S, rsave, true);
    double f[3];
    for (int c = 0; c < 3; ++c)
        f[c] = sizeof(R) == 4 ? (double)A.f[(size_t)(F_FX + c) * n + i] : A.d[(size_t)(D_FX + c) * n + i];
    langevin_add(T, call, orig[i], pi.x, fin.v, fin.m, f);
    finalize_atom<R>(i, pi, f[0], f[1], f[2], A, S, pos_next, fin, desc[w_sys(pi.w)], misc);
}

static void langevin_apply(Batch& b, int prec) {
    if (b.therm.count == 0 || b.live == 0) return;
    AtomArrays AA = arrays(b);
    hipLaunchKernelGGL(k_langevin_apply, dim3(nblk(b.live, 256)), dim3(256), 0, b.ctx->stream, b.pos[b.cur], AA,
                       b.orig[b.sc], b.therm, b.therm_calls, prec);
    b.therm_calls++;
}

void stream_wait(Ctx& c) {
    if (!c.spin) {
        HIPCHECK(hipStreamSynchronize(c.stream));
        return;
    }
    hipError_t e;
    while ((e = hipStreamQuery(c.stream)) == hipErrorNotReady) {
    }
    HIPCHECK(e);
}

void event_wait(Ctx& c, hipEvent_t ev) {
    if (!c.spin) {
        HIPCHECK(hipEventSynchronize(ev));
        return;
    }
    hipError_t e;
    while ((e = hipEventQuery(ev)) == hipErrorNotReady) {
    }
    HIPCHECK(e);
}

__global__ void k_zero(double* __restrict__ p, long long count) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0.0;
}

// The device work of one evaluation: zero what the evaluation does not
// write, the force (and density) kernels, the counters to the host slot.
// Fill kernels rather than memsets, so that the sequence can be captured.
static void compute_sequence(Batch& b, const Pot& p, int what, int prec) {
    hipStream_t st = b.ctx->stream;
    StepArgs S{0.0, -1, MODE_COMPUTE, what & (W_ENERGY | W_VIRIAL), 1};
    // the reference's prepare zeroes Ep, Em and V and only virials() fills
    // them (solver/single.cxx:38-53)
    if (p.type != PT_BPNN && !(what & W_ENERGY))
        hipLaunchKernelGGL(k_zero, dim3(nblk(2LL * b.n, 256)), dim3(256), 0, st, b.d(b.sc, D_EP), 2LL * b.n);
    if (!(what & W_VIRIAL))
        hipLaunchKernelGGL(k_zero, dim3(nblk(6LL * b.n, 256)), dim3(256), 0, st, b.d(b.sc, D_VXX), 6LL * b.n);
    launch_step_kernels(b, p, S, prec, b.pos[b.cur], nullptr);
    langevin_apply(b, prec);
    HIPCHECK(hipMemcpyAsync(b.ctx->h_pin, b.misc, sizeof(int) * M_COUNT, hipMemcpyDeviceToHost, st));
}

void graph_release(Batch& b) {
    if (b.gexec) (void)hipGraphExecDestroy(b.gexec);
    b.gexec = nullptr;
    b.gkey_set = false;
}

// Graph mode: replay the captured evaluation of this configuration, or
// capture it if the last call ran the same configuration eagerly, or run a
// new configuration eagerly and remember it.
static void compute_graph(Batch& b, const Pot& p, int what, int prec) {
    Batch::GraphKey k;
    k.pot = &p;
    k.what = what;
    k.prec = prec;
    k.rebuilds = b.rebuilds;
    k.cap = b.cap;
    k.live = b.live;
    k.tpa = b.tpa;
    k.block = b.block;
    k.cur = b.cur;
    k.sc = b.sc;
    k.alloc_gen = b.alloc_gen;
    hipStream_t st = b.ctx->stream;
    if (b.gkey_set && b.gkey == k) {
        if (!b.gexec) {
            hipGraph_t g = nullptr;
            HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                compute_sequence(b, p, what, prec);
            } catch (...) {
                (void)hipStreamEndCapture(st, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIPCHECK(hipStreamEndCapture(st, &g));
            const hipError_t e = hipGraphInstantiate(&b.gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                b.gexec = nullptr;
                throw Error(E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
            }
            b.graph_captures++;
        }
        HIPCHECK(hipGraphLaunch(b.gexec, st));
        b.graph_launches++;
        return;
    }
    // a new configuration: drop the old graph; this call runs eagerly and
    // its configuration (after any allocation it makes) is the next key
    graph_release(b);
    compute_sequence(b, p, what, prec);
    k.alloc_gen = b.alloc_gen;
    k.cap = b.cap;
    b.gkey = k;
    b.gkey_set = true;
}

void force_compute(Batch& b, const Pot& p, int what, int prec) {
    if (b.n == 0) return;
    set_trigger(b, p.rcut, false);
    if (!b.list_valid) nlist_rebuild(b);
    // graphs hold the thermostat's random-stream counter and the arrays
    // they were captured with: eager with thermostats, timing or a borrowed
    // work area
    if (b.graphs && b.therm.count == 0 && !b.timing && !b.bp_fixed) compute_graph(b, p, what, prec);
    else compute_sequence(b, p, what, prec);
    b.f_prec = prec;
    b.last_rcut = p.rcut;
    stream_wait(*b.ctx);
    timer_drain(b);
    check_errors(b);
}

// Per-atom prepare of solver::Single and solver::Ternary
// (single.cxx:38-53, ternary.cxx:26-40): atoms with tag != 0 get zero
// force, virial and pair energy, and for EAM zero embedding energy and
// density.
__global__ void k_prepare(const double4* __restrict__ pos, AtomArrays A, int eam) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = A.n;
    if (i >= A.live || w_inactive(pos[i].w)) return;
    for (int f = D_FX; f <= D_FZ; ++f) A.d[(size_t)f * n + i] = 0.0;
    for (int f = F_FX; f <= F_FZ; ++f) A.f[(size_t)f * n + i] = 0.0f;
    for (int f = D_VXX; f <= D_VYZ; ++f) A.d[(size_t)f * n + i] = 0.0;
    A.d[(size_t)D_EP * n + i] = 0.0;
    if (eam) {
        A.d[(size_t)D_EM * n + i] = 0.0;
        A.d[(size_t)D_NC * n + i] = 0.0;
    }
}

// virials() adds the pair energy and virial to what the atoms hold and
// leaves their force (pair.cxx:100-127, LJs-cut.cxx:86-131,
// EAM.cxx:306-354).  The kernels wrote fresh values into the current
// state; `was` holds the state from before the call.
__global__ void k_virial_merge(const double4* __restrict__ pos, AtomArrays A, const double* __restrict__ was_d,
                               const float* __restrict__ was_f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = A.n;
    if (i >= A.live || w_inactive(pos[i].w)) return;
    for (int f = D_FX; f <= D_FZ; ++f) A.d[(size_t)f * n + i] = was_d[(size_t)f * n + i];
    for (int f = F_FX; f <= F_FZ; ++f) A.f[(size_t)f * n + i] = was_f[(size_t)f * n + i];
    A.d[(size_t)D_EP * n + i] += was_d[(size_t)D_EP * n + i];
    for (int f = D_VXX; f <= D_VYZ; ++f) A.d[(size_t)f * n + i] += was_d[(size_t)f * n + i];
}

// forces() adds to the force the atoms hold (single.cxx:65-73,
// pair.cxx:72-87); the kernels wrote the fresh force of precision `prec`,
// `was` holds the state from before the call with its force in `was_prec`.
__global__ void k_force_merge(const double4* __restrict__ pos, AtomArrays A, const double* __restrict__ was_d,
                              const float* __restrict__ was_f, int was_prec, int prec) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = A.n;
    if (i >= A.live || w_inactive(pos[i].w)) return;
    for (int c = 0; c < 3; ++c) {
        const double w = was_prec == P_MIXED ? (double)was_f[(size_t)(F_FX + c) * n + i]
                                             : was_d[(size_t)(D_FX + c) * n + i];
        if (prec == P_MIXED) A.f[(size_t)(F_FX + c) * n + i] = (float)(w + (double)A.f[(size_t)(F_FX + c) * n + i]);
        else A.d[(size_t)(D_FX + c) * n + i] += w;
    }
}

// One solver call of the reference's Python API on the resident state:
// prepare, charge, forces or virials (pair.cxx:16-127, single.cxx:23-83,
// ternary.cxx:19-67).  Each reads what the atoms hold, as the reference's
// do: forces() without a charge() before it uses the stored density.
void stage_compute(Batch& b, const Pot& p, int stage, int prec) {
    if (b.n == 0) return;
    if (!b.list_valid) throw Error(E_STATE, "no neighbour list: update the list first");
    hipStream_t st = b.ctx->stream;
    set_trigger(b, p.rcut, false);
    const bool eam = p.type == PT_EAM;
    const bool bpnn = p.type == PT_BPNN;
    const int grid = nblk(b.live, 256);
    AtomArrays AA = arrays(b);
    const double4* pos = b.pos[b.cur];
    auto pass = [&](int which, const StepArgs& S) {
        if (prec == P_MIXED) launch_pair_r<float>(b, p, S, pos, nullptr, which);
        else launch_pair_r<double>(b, p, S, pos, nullptr, which);
    };
    switch (stage) {
        case ST_PREPARE:
            if (grid) hipLaunchKernelGGL(k_prepare, dim3(grid), dim3(256), 0, st, pos, AA, eam ? 1 : 0);
            // solver::Multi's prepare leaves rcut (multi.cxx:56-66)
            if (!(eam && p.eam.nk > 1 && !p.eam.alloy)) b.last_rcut = p.rcut;
            b.f_prec = P_FP64;
            b.bp_prepared = -1;
            if (bpnn) {
                // the network runs in prepare (maise_bpnn.cxx:27-59, :275-325)
                bpnn_compute(b, p, W_ENERGY, prec, MODE_COMPUTE, 0.0, -1, pos, nullptr, 1);
                b.bp_prepared = prec;
                b.bp_prepared_builds = b.rebuilds;
            }
            break;
        case ST_CHARGE:
            if (eam) pass(1, StepArgs{0.0, -1, MODE_COMPUTE, 0, 1});
            break;
        case ST_FORCES: {
            const int idle = b.sc ^ 1;
            const int was_prec = b.f_prec < 0 ? P_FP64 : b.f_prec;
            if (b.f_prec < 0) HIPCHECK(hipMemsetAsync(b.d(idle, D_FX), 0, sizeof(double) * 3 * b.n, st));
            else if (was_prec == P_MIXED)
                HIPCHECK(hipMemcpyAsync(b.fl(idle, F_FX), b.fl(b.sc, F_FX), sizeof(float) * 3 * b.n,
                                        hipMemcpyDeviceToDevice, st));
            else
                HIPCHECK(hipMemcpyAsync(b.d(idle, D_FX), b.d(b.sc, D_FX), sizeof(double) * 3 * b.n,
                                        hipMemcpyDeviceToDevice, st));
            if (bpnn) {
                if (b.bp_prepared != prec || b.bp_prepared_builds != b.rebuilds)
                    throw Error(E_STATE, "BPNN forces need a prepare at the same precision on the current list");
                bpnn_compute(b, p, 0, prec, MODE_COMPUTE, 0.0, -1, pos, nullptr, 2);
            } else {
                if (eam) pass(1, StepArgs{0.0, -1, MODE_COMPUTE, W_NC_STORED, 1});
                pass(2, StepArgs{0.0, -1, MODE_COMPUTE, 0, 1});
            }
            if (grid)
                hipLaunchKernelGGL(k_force_merge, dim3(grid), dim3(256), 0, st, pos, AA, b.dst[idle], b.fst[idle],
                                   was_prec, prec);
            langevin_apply(b, prec);
            b.f_prec = prec;
            break;
        }
        case ST_VIRIALS: {
            if (bpnn) {
                // the diagonal virial of maise_bpnn.cxx:98-120 from the
                // dE/dG of the last prepare, added to V
                if (b.bp_prepared != prec || b.bp_prepared_builds != b.rebuilds)
                    throw Error(E_STATE, "BPNN virials need a prepare at the same precision on the current list");
                bpnn_compute(b, p, 0, prec, MODE_COMPUTE, 0.0, -1, pos, nullptr, 2 | 4);
                break;
            }
            const int idle = b.sc ^ 1;
            HIPCHECK(hipMemcpyAsync(b.dst[idle], b.dst[b.sc], sizeof(double) * D_NFIELDS * b.n,
                                    hipMemcpyDeviceToDevice, st));
            HIPCHECK(hipMemcpyAsync(b.fst[idle], b.fst[b.sc], sizeof(float) * F_NFIELDS * b.n,
                                    hipMemcpyDeviceToDevice, st));
            if (eam) pass(1, StepArgs{0.0, -1, MODE_COMPUTE, W_NC_STORED | W_ENERGY, 1});
            pass(2, StepArgs{0.0, -1, MODE_COMPUTE, W_ENERGY | W_VIRIAL, 1});
            if (grid) hipLaunchKernelGGL(k_virial_merge, dim3(grid), dim3(256), 0, st, pos, AA, b.dst[idle], b.fst[idle]);
            break;
        }
        default:
            throw Error(E_INVAL, "stage must be prepare (0), charge (1), forces (2) or virials (3)");
    }
    read_misc(b);
    check_errors(b);
}

// The fixed-point step path applies to single-element EAM in mixed
// precision without thermostats, with every box's minimum image unique
// between rebuilds (set_trigger has run); the box index must fit the 22
// bits the record's tag word keeps.
static bool qfast_ok(const Batch& b, const Pot& p, int prec) {
    return prec == P_MIXED && p.type == PT_EAM && p.eam.nk == 1 && !p.eam.alloy && p.eam.r32p &&
           b.therm.count == 0 && b.all_shift && !b.manual_rebuild && b.live > 0 && b.nsys < (1 << 22) && !b.q_off;
}

static void qfast_alloc(Batch& b) {
    if (b.uq_n >= b.n && b.uq[0]) return;
    for (uint4** a : {&b.uq[0], &b.uq[1], &b.uqF}) {
        if (*a) HIPCHECK(hipFree(*a));
        *a = nullptr;
    }
    b.uq_n = 0;
    for (uint4** a : {&b.uq[0], &b.uq[1], &b.uqF})
        if (hipMalloc(a, sizeof(uint4) * std::max(b.n, 1)) != hipSuccess)
            throw Error(E_NOMEM, "out of device memory for the fixed-point positions");
    b.uq_n = b.n;
}

// fixed-point records of the current positions (after a sort)
static void qfast_refresh(Batch& b) {
    hipLaunchKernelGGL(k_fixq, dim3(nblk(b.live, 256)), dim3(256), 0, b.ctx->stream, b.pos[b.cur], b.uq[b.cur],
                       b.d_desc, b.live);
}

void run_steps(Batch& b, const Pot& p, double dt, int nsteps, int prec, int chunk) {
    if (b.n == 0 || nsteps <= 0) return;
    hipStream_t st = b.ctx->stream;
    if (!b.list_valid) nlist_rebuild(b);
    set_trigger(b, p.rcut, true);
    if (b.f_prec < 0) throw Error(E_STATE, "no force to start from: compute forces once before stepping");
    if (b.live == 0) return;  // every record void: nothing moves
    const int grid = nblk(b.live, 256);
    int abort_init = INT_MAX;
    HIPCHECK(hipMemcpyAsync(b.misc + M_ABORT, &abort_init, sizeof(int), hipMemcpyHostToDevice, st));
    const bool qf = qfast_ok(b, p, prec);
    if (qf) qfast_alloc(b);
    struct QOff {
        Batch& b;
        ~QOff() { b.qfast = false; }
    } qoff{b};
    b.qfast = qf;
    b.q_last = qf ? (b.q_lds ? 2 : 1) : 0;
    int c0 = b.cur;
    {
        AtomArrays AA = arrays(b);
        launch_k(b, K_PREDRIFT, k_predrift, dim3(grid), dim3(256), 0, b.pos[c0], b.pos[c0 ^ 1], AA, b.f_prec, dt,
                 b.d_desc, b.rsave, b.misc, b.qfast ? b.uq[c0 ^ 1] : (uint4*)nullptr);
    }
    auto buf = [&](int t) { return c0 ^ 1 ^ (t & 1); };
    // Steps go out in chunks, two chunks in flight: while one runs, the
    // host reads the trigger and error counters the other left (copied to
    // a pinned slot behind it), so the device never waits for the host
    // between chunks.  Kernels after a raised trigger exit at once.
    struct Chunk {
        int s, e, slot;
    };
    Chunk q[2];
    int nq = 0, head = 0, next_slot = 0;
    int s = 0, launch_s = 0;
    auto launch_chunk = [&](int a, int e) {
        for (int t = a; t < e; ++t) {
            const bool last = t == nsteps - 1;
            StepArgs S{dt, t, last ? MODE_FINAL : MODE_FINAL_DRIFT, 0, last ? 1 : 0};
            if (b.therm.count == 0) {
                launch_step_kernels(b, p, S, prec, b.pos[buf(t)], last ? nullptr : b.pos[buf(t) ^ 1]);
                continue;
            }
            StepArgs SF = S;
            SF.mode = MODE_FORCE_STEP;
            launch_step_kernels(b, p, SF, prec, b.pos[buf(t)], nullptr);
            AtomArrays AA = arrays(b);
            auto fin = prec == P_MIXED ? k_langevin_finish<float> : k_langevin_finish<double>;
            hipLaunchKernelGGL(fin, dim3(grid), dim3(256), 0, st, b.pos[buf(t)], last ? nullptr : b.pos[buf(t) ^ 1],
                               AA, S, b.orig[b.sc], b.therm, b.therm_calls + (uint64_t)t, b.d_desc, b.rsave, b.misc);
        }
        int* slot = b.ctx->h_pin + M_COUNT * (1 + next_slot);
        HIPCHECK(hipMemcpyAsync(slot, b.misc, sizeof(int) * M_COUNT, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipEventRecord(b.ctx->chunk_ev[next_slot], st));
        q[(head + nq) % 2] = Chunk{a, e, next_slot};
        ++nq;
        next_slot ^= 1;
    };
    while (s < nsteps) {
        while (nq < 2 && launch_s < nsteps) {
            // up to the predicted next rebuild in one chunk, then short ones
            int c = chunk;
            if (b.pred_interval > 0) {
                const int left = (int)(0.8 * b.pred_interval) - b.since_rebuild - (launch_s - s);
                c = left > 2 ? left : 2;
            }
            c = std::max(1, std::min(c, 64));
            const int e = std::min(nsteps, launch_s + c);
            launch_chunk(launch_s, e);
            launch_s = e;
        }
        const Chunk P = q[head];
        head = (head + 1) % 2;
        --nq;
        event_wait(*b.ctx, b.ctx->chunk_ev[P.slot]);
        const int* cnt = b.ctx->h_pin + M_COUNT * (1 + P.slot);
        check_errors(b, cnt);
        const int ab = cnt[M_ABORT];
        if (cnt[M_MAXNN] > b.cap) {
            // the rebuild before this chunk overflowed the list, and the
            // abort flag it raised stopped every step since (s = P.s): grow
            // the list, build it again and go on from there
            if (nq) {
                HIPCHECK(hipStreamSynchronize(st));
                nq = 0;
                head = 0;
            }
            nlist_regrow(b, cnt[M_MAXNN]);
            launch_s = s;
            continue;
        }
        if (ab <= P.e) {
            // steps [s, ab) ran; positions of step ab wait in buf(ab) (the
            // last step of a sequence never raises the trigger, so ab <
            // nsteps).  A chunk queued behind this one exits at once; the
            // rebuild queues behind it without waiting, and its counters
            // and event are never read.
            nq = 0;
            head = 0;
            b.since_rebuild += ab - s;
            b.pred_interval = b.pred_interval > 0 ? 0.5 * b.pred_interval + 0.5 * b.since_rebuild
                                                  : (double)b.since_rebuild;
            b.since_rebuild = 0;
            b.cur = buf(ab);
            nlist_rebuild(b, false);  // with b.qfast it also writes the sorted records
            timer_drain(b);
            c0 = b.cur ^ 1 ^ (ab & 1);
            s = launch_s = ab;  // the rebuild cleared the trigger
        } else {
            b.since_rebuild += P.e - P.s;
            s = P.e;
        }
    }
    timer_drain(b);
    b.cur = buf(nsteps - 1);
    if (b.therm.count) b.therm_calls += (uint64_t)nsteps;
    // after make_step the reference holds Ep = Em = 0 and V = 0 (prepare
    // zeroes them; only virials() fills them); BPNN sets Ep in prepare
    if (p.type != PT_BPNN) {
        HIPCHECK(hipMemsetAsync(b.d(b.sc, D_EP), 0, sizeof(double) * 2 * b.n, st));
    }
    HIPCHECK(hipMemsetAsync(b.d(b.sc, D_VXX), 0, sizeof(double) * 6 * b.n, st));
    b.f_prec = prec;
    b.last_rcut = p.rcut;
}

// Mean duration of the step-mode density and force kernels (the
// configuration of every step but the last in a sequence), each launched
// reps times back to back between two events on the context stream.  The
// velocities, which the force kernel updates in place, forces and F' are
// saved in the idle half of the double-buffered state and restored.
void time_step_kernels(Batch& b, const Pot& p, int prec, double dt, int reps, double* us) {
    us[0] = us[1] = 0.0;
    if (b.n == 0) return;
    if (p.type == PT_BPNN) throw Error(E_INVAL, "kernel timing covers the LJ and EAM kernels");
    if (b.f_prec < 0) throw Error(E_STATE, "no force to start from: compute forces once before timing");
    hipStream_t st = b.ctx->stream;
    if (!b.list_valid) nlist_rebuild(b);
    if (b.live == 0) return;
    set_trigger(b, p.rcut, true);
    const int idle = b.sc ^ 1;
    const size_t vbytes = sizeof(double) * 3 * (size_t)b.n;
    HIPCHECK(hipMemcpyAsync(b.d(idle, D_VX), b.d(b.sc, D_VX), vbytes, hipMemcpyDeviceToDevice, st));
    // forces and F' too: the kernels a step sequence runs may round them
    // differently from the evaluation that stored them
    const size_t fbytes = sizeof(float) * 3 * (size_t)b.n;
    HIPCHECK(hipMemcpyAsync(b.fl(idle, F_FX), b.fl(b.sc, F_FX), fbytes, hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipMemcpyAsync(b.d(idle, D_FX), b.d(b.sc, D_FX), vbytes, hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipMemcpyAsync(b.d(idle, D_FP), b.d(b.sc, D_FP), sizeof(double) * (size_t)b.n,
                            hipMemcpyDeviceToDevice, st));
    int abort_init = INT_MAX;
    HIPCHECK(hipMemcpyAsync(b.misc + M_ABORT, &abort_init, sizeof(int), hipMemcpyHostToDevice, st));
    const StepArgs S{dt, 0, MODE_FINAL_DRIFT, 0, 0};
    const double4* pos = b.pos[b.cur];
    double4* pos_next = b.pos[b.cur ^ 1];
    // the kernels a step sequence of this configuration would run
    const bool qf = qfast_ok(b, p, prec);
    if (qf) {
        qfast_alloc(b);
        qfast_refresh(b);
    }
    struct QOff {
        Batch& b;
        ~QOff() { b.qfast = false; }
    } qoff{b};
    b.qfast = qf;
    b.q_last = qf ? (b.q_lds ? 2 : 1) : 0;
    hipEvent_t e[3];
    for (auto& x : e) HIPCHECK(hipEventCreate(&x));
    const bool timing = b.timing;
    b.timing = false;
    auto run = [&](int which) {
        if (prec == P_MIXED) launch_pair_r<float>(b, p, S, pos, pos_next, which);
        else launch_pair_r<double>(b, p, S, pos, pos_next, which);
    };
    try {
        run(3);  // warm-up
        HIPCHECK(hipEventRecord(e[0], st));
        if (p.type == PT_EAM)
            for (int r = 0; r < reps; ++r) run(1);
        HIPCHECK(hipEventRecord(e[1], st));
        for (int r = 0; r < reps; ++r) run(2);
        HIPCHECK(hipEventRecord(e[2], st));
        HIPCHECK(hipEventSynchronize(e[2]));
    } catch (...) {
        b.timing = timing;
        for (auto& x : e) (void)hipEventDestroy(x);
        throw;
    }
    b.timing = timing;
    float m0 = 0.f, m1 = 0.f;
    HIPCHECK(hipEventElapsedTime(&m0, e[0], e[1]));
    HIPCHECK(hipEventElapsedTime(&m1, e[1], e[2]));
    for (auto& x : e) HIPCHECK(hipEventDestroy(x));
    us[0] = p.type == PT_EAM ? 1e3 * m0 / reps : 0.0;
    us[1] = 1e3 * m1 / reps;
    HIPCHECK(hipMemcpyAsync(b.d(b.sc, D_VX), b.d(idle, D_VX), vbytes, hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipMemcpyAsync(b.fl(b.sc, F_FX), b.fl(idle, F_FX), fbytes, hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipMemcpyAsync(b.d(b.sc, D_FX), b.d(idle, D_FX), vbytes, hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipMemcpyAsync(b.d(b.sc, D_FP), b.d(idle, D_FP), sizeof(double) * (size_t)b.n,
                            hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipMemcpyAsync(b.misc + M_ABORT, &abort_init, sizeof(int), hipMemcpyHostToDevice, st));
    read_misc(b);
    check_errors(b);
}

void thermo(Batch& b, double* out) {
    hipStream_t st = b.ctx->stream;
    // each box's atoms: the first present[s] of its records in the
    // caller's order (the rest are void)
    std::vector<int> off(2 * (size_t)b.nsys);
    for (int s = 0; s < b.nsys; ++s) {
        off[2 * s] = b.atom_off[s];
        off[2 * s + 1] = b.present[s];
    }
    int* d_off = reinterpret_cast<int*>(b.thermo + 8 * b.nsys);
    HIPCHECK(hipMemcpyAsync(d_off, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_thermo, dim3(b.nsys), dim3(256), 0, st, b.dst[b.sc], b.n, d_off, b.where, b.thermo);
    HIPCHECK(hipMemcpyAsync(out, b.thermo, sizeof(double) * 8 * b.nsys, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
}

}  // namespace mdc
