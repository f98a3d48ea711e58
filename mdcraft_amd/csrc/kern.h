// Device helpers shared by the pair, EAM and BPNN kernels: step arguments,
// the abort check of a step sequence and the fused velocity-Verlet finish.
#pragma once
#include "engine.h"

namespace mdc {

// MODE_FORCE_STEP: a force pass inside a step sequence that only stores
// the force; a separate kernel adds the thermostat force and integrates
enum Mode { MODE_COMPUTE = 0, MODE_FINAL = 1, MODE_FINAL_DRIFT = 2, MODE_FORCE_STEP = 3 };

__device__ inline float rcp_(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline double rcp_(double x) { return 1.0 / x; }

template <typename R>
struct Vec4T;
template <>
struct Vec4T<double> { using T = double4; };
template <>
struct Vec4T<float> { using T = float4; };

// Hermite interval of the reference tables (EAM.cxx:367-394).
// FP64 records are (y0, z0, y1, z1) and use the reference's formula.  FP32
// records hold the same cubic as coefficients in the interval fraction t,
// c0 + t (c1 + t (c2 + t c3)), computed in FP64 on the host: forming the
// slope (y1 - y0)/dr from FP32 knots cancels catastrophically.
template <typename R>
__device__ inline void spline_vd(const typename Vec4T<R>::T rec, R R1, R RN, int N1, R dr, R dr1, R& val, R& der);
template <typename R>
__device__ inline R spline_v(const typename Vec4T<R>::T rec, R R1, R RN, int N1, R dr, R dr1);

template <>
__device__ inline void spline_vd<double>(const double4 rec, double R1, double RN, int N1, double dr, double dr1,
                                         double& val, double& der) {
    double VR = (double)N1;
    double DV = RN - VR;
    double h = R1 - VR * dr;
    double s = (rec.z - rec.x) * dr1;
    double aa = (rec.w - s) - (s - rec.y);
    double bb = aa - (s - rec.y);
    val = rec.x + h * (rec.y + DV * (DV * aa - bb));
    der = rec.y + 2.0 * DV * (1.5 * DV * aa - bb);
}

template <>
__device__ inline double spline_v<double>(const double4 rec, double R1, double RN, int N1, double dr, double dr1) {
    double VR = (double)N1;
    double DV = RN - VR;
    double h = R1 - VR * dr;
    double s = (rec.z - rec.x) * dr1;
    double aa = (rec.w - s) - (s - rec.y);
    double bb = aa - (s - rec.y);
    return rec.x + h * (rec.y + DV * (DV * aa - bb));
}

template <>
__device__ inline void spline_vd<float>(const float4 c, float R1, float RN, int N1, float dr, float dr1, float& val,
                                        float& der) {
    const float t = RN - (float)N1;
    val = c.x + t * (c.y + t * (c.z + t * c.w));
    der = (c.y + t * (2.f * c.z + 3.f * t * c.w)) * dr1;
}

template <>
__device__ inline float spline_v<float>(const float4 c, float R1, float RN, int N1, float dr, float dr1) {
    const float t = RN - (float)N1;
    return c.x + t * (c.y + t * (c.z + t * c.w));
}

struct StepArgs {
    double dt;
    int step;  // index of this step inside the running sequence (-1: plain compute)
    int mode;
    int flags;  // W_ENERGY | W_VIRIAL
    int store_aux;
};

struct AtomArrays {
    double* d;   // current double fields base (stride n)
    float* f;    // current float fields base (stride n)
    int n;
    int live;    // records [0, live) hold atoms; the rest are void
};

inline AtomArrays arrays(const Batch& b) { return AtomArrays{b.dst[b.sc], b.fst[b.sc], b.n, b.live}; }

// Per-atom inputs of the velocity-Verlet finish, loaded in the kernel
// prologue so their latency overlaps the neighbour loop.
struct FinIn {
    double v[3];
    double m;
    double4 rs;
};

__device__ inline FinIn load_fin(int i, const AtomArrays& A, const StepArgs& S, const double4* __restrict__ rsave,
                                 bool need) {
    FinIn f;
    if (need && (S.mode == MODE_FINAL || S.mode == MODE_FINAL_DRIFT)) {
        const int n = A.n;
        f.v[0] = A.d[(size_t)D_VX * n + i];
        f.v[1] = A.d[(size_t)D_VY * n + i];
        f.v[2] = A.d[(size_t)D_VZ * n + i];
        f.m = A.d[(size_t)D_M * n + i];
        f.rs = S.mode == MODE_FINAL_DRIFT ? rsave[i] : make_double4(0, 0, 0, 0);
    } else {
        f.v[0] = f.v[1] = f.v[2] = 0.0;
        f.m = 1.0;
        f.rs = make_double4(0, 0, 0, 0);
    }
    return f;
}

// Finalise one atom after its force is known: store force, and for step
// modes apply the closing half kick (verlet.cxx:80-85) and, when another
// step follows, the opening half kick and drift (verlet.cxx:71-78) into the
// other position buffer together with the displacement trigger of
// verlet-list.cxx:149-154.
// uq_next (fixed-point path, else null): the drifted position's
// fixed-point record as well.
template <typename R>
__device__ inline void finalize_atom(int i, const double4& pi, double fx, double fy, double fz, const AtomArrays& A,
                                     const StepArgs& S, double4* __restrict__ pos_next, const FinIn& in,
                                     const SysDesc& s, int* __restrict__ misc, uint4* __restrict__ uq_next = nullptr) {
    const int n = A.n;
    if (sizeof(R) == 4) {
        A.f[(size_t)F_FX * n + i] = (float)fx;
        A.f[(size_t)F_FY * n + i] = (float)fy;
        A.f[(size_t)F_FZ * n + i] = (float)fz;
    } else {
        A.d[(size_t)D_FX * n + i] = fx;
        A.d[(size_t)D_FY * n + i] = fy;
        A.d[(size_t)D_FZ * n + i] = fz;
    }
    if (S.mode != MODE_FINAL && S.mode != MODE_FINAL_DRIFT) return;
    const double m = in.m;
    double v[3] = {in.v[0], in.v[1], in.v[2]};
    const double F[3] = {fx, fy, fz};
    for (int c = 0; c < 3; ++c) {
        double a = F[c] / m;
        v[c] += a * 0.5 * S.dt;
    }
    if (S.mode == MODE_FINAL_DRIFT) {
        double r[3] = {pi.x, pi.y, pi.z};
        for (int c = 0; c < 3; ++c) {
            double a = F[c] / m;
            v[c] = v[c] + a * S.dt * 0.5;
            r[c] += v[c] * S.dt;
        }
        pos_next[i] = make_double4(r[0], r[1], r[2], pi.w);
        if (uq_next) uq_next[i] = fixq4(r[0], r[1], r[2], pi.w, s);
        const double4 q = in.rs;
        double dx = r[0] - q.x, dy = r[1] - q.y, dz = r[2] - q.z;
        shortest3(s, dx, dy, dz);
        wave_atomic_min(sq3_exact(dx, dy, dz) > s.trig2, &misc[M_ABORT], S.step + 1);
        wave_atomic_or(!(fabs(r[0]) < 1e300 && fabs(r[1]) < 1e300 && fabs(r[2]) < 1e300), &misc[M_ERR],
                       EB_NONFINITE);
    }
    A.d[(size_t)D_VX * n + i] = v[0];
    A.d[(size_t)D_VY * n + i] = v[1];
    A.d[(size_t)D_VZ * n + i] = v[2];
}

// The abort flag of a step sequence, read without ordering so the load
// overlaps the other prologue loads.
__device__ inline int abort_flag(const StepArgs& S, const int* misc) {
    return S.mode != MODE_COMPUTE ? __atomic_load_n(&misc[M_ABORT], __ATOMIC_RELAXED) : 0x7fffffff;
}

// Neighbour slots of this lane are visited in groups of U: the U list
// reads, then the U position gathers (32-bit-offset buffer loads), are
// issued back to back so their latencies overlap.  j0 holds the first
// group's list entries, read in the kernel prologue (the list stores M chunks
// per atom, so those reads stay inside it).
template <int TPA, int U>
__device__ inline void first_group(const AtomSlots& as, bool valid, int M, const int* __restrict__ nl, int* j0) {
#pragma unroll
    for (int u = 0; u < U; ++u) j0[u] = (valid && u < M) ? nl[as.base + (long long)u * kWave] : 0;
}

// The entries of the group after the one starting at chunk m0, read while
// the current group's gathers are in flight (chunks past M are not read).
template <int TPA, int U>
__device__ inline void next_group(const AtomSlots& as, int m0, int M, const int* __restrict__ nl, int* en) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int c = m0 + U + u;
        en[u] = c < M ? nl[as.base + (long long)c * kWave] : 0;
    }
}

// The same with every read issued (chunks past M read the last chunk, whose
// entries the loop then ignores), so that the compiler can count the loads
// in flight and wait for the current group's gathers alone.
template <int U>
__device__ inline void next_group_clamped(const AtomSlots& as, int m0, int M, const int* __restrict__ nl, int* en) {
#pragma unroll
    for (int u = 0; u < U; ++u) en[u] = nl[as.base + (long long)min(m0 + U + u, M - 1) * kWave];
}

// Difference r_j - r_i of a listed pair under the reference's image rule
// (pair_image); when the system's boxes make the minimum image unique
// between rebuilds, the shift recorded at build time is applied instead.
__device__ inline bool pair_delta(const Frame& s, const double4& pi, const double4& pj, int e, double& dx,
                                  double& dy, double& dz, double& d2) {
    dx = pj.x - pi.x;
    dy = pj.y - pi.y;
    dz = pj.z - pi.z;
    if (s.shift_ok) {
        dx += (double)ent_s(e, 0) * s.L[0];
        dy += (double)ent_s(e, 1) * s.L[1];
        dz += (double)ent_s(e, 2) * s.L[2];
        d2 = sq3_exact(dx, dy, dz);
        return true;
    }
    return pair_image(s, dx, dy, dz, d2);
}

// Chunks of the longest list among the wave's lanes, the same value in
// every lane: neighbour loops over it run all lanes the same number of
// times, as the cooperative 32-byte loads (load32_pair) need.
template <int TPA>
__device__ inline int wave_chunks(int cnt) {
    int c = (cnt + TPA - 1) / TPA;
    for (int o = kWave / 2; o > 0; o >>= 1) c = max(c, __shfl_xor(c, o, kWave));
    return __builtin_amdgcn_readfirstlane(c);
}

template <int TPA, int U, class Body>
__device__ inline void for_neighbours(const AtomSlots& as, int cnt, const int* __restrict__ nl,
                                      __amdgpu_buffer_rsrc_t posb, const int* j0, Body&& body) {
    const int nch = wave_chunks<TPA>(cnt);
    for (int m0 = 0; m0 < nch; m0 += U) {
        int e[U];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = (m0 + u) * TPA + as.sub;
            ok[u] = k < cnt;
            e[u] = ok[u] ? (m0 == 0 ? j0[u] : nl[as.base + (long long)(m0 + u) * kWave]) : 0;
        }
        double4 pj[U];
#pragma unroll
        for (int u = 0; u < U; ++u) pj[u] = buf_d4_pair(posb, ok[u] ? (unsigned)ent_j(e[u]) * 32u : kSkip);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (ok[u]) body(e[u], pj[u]);
    }
}

__device__ inline bool skip_step(const StepArgs& S, const int* misc) {
    return S.mode != MODE_COMPUTE && *(volatile const int*)&misc[M_ABORT] <= S.step;
}

}  // namespace mdc
