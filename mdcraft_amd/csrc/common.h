// Shared device-side definitions of the MI355X force-and-integration engine.
//
// Layout in HBM (one "batch" = one or more independent periodic/open boxes
// packed back to back; a single system is a batch of one):
//   pos[2]   double4 per atom: x, y, z and a tag word (kind | inactive<<8 |
//            system<<9, stored as an exact integer in the double); two
//            buffers so a fused kernel can drift into the other one while
//            neighbours are still read from the first.
//   vel      double, SoA [3][N]
//   frc      Real (double or float), SoA [3][N]
//   mass     double [N]; rsave double4 [N] (positions at the last build)
//   list     int32, wavefront-packed: for wave w (A = 64/TPA atoms, TPA
//            lanes per atom), chunk m and lane l the entry is
//            list[(w*M + m)*64 + l] with slot k = m*TPA + l%TPA of atom
//            w*A + l/TPA; every load instruction of a wave reads 256
//            contiguous bytes.  Self is not stored.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mdc {

constexpr int kWave = 64;
constexpr int kMaxKinds = 8;

struct SysDesc {
    double lo[3], hi[3], L[3];
    double csize[3];
    double rns2;      // (max rns)^2: image rule of the ghost pass and cell bound
    double trig2;     // 0.25 * skin^2 for the displacement trigger (set per step)
    int per[3];
    int nc[3];
    int cell_off;
    int atom_off;
    int natoms;
    int any_per;
};

// tag word packed into pos.w
__host__ __device__ inline double pack_w(int kind, int inactive, int sys) {
    return (double)((long long)kind | ((long long)inactive << 8) | ((long long)sys << 9));
}
__device__ inline int w_kind(double w) { return (int)((long long)w & 0xff); }
__device__ inline int w_inactive(double w) { return (int)(((long long)w >> 8) & 1); }
__device__ inline int w_sys(double w) { return (int)((long long)w >> 9); }

// lattice::Domain::shortest (mdcraft/lattice/domain.cxx:59-66)
__device__ inline void shortest3(const SysDesc& s, double& x, double& y, double& z) {
    if (s.per[0]) {
        x = fabs(x + s.L[0]) < fabs(x) ? x + s.L[0] : x;
        x = fabs(x - s.L[0]) < fabs(x) ? x - s.L[0] : x;
    }
    if (s.per[1]) {
        y = fabs(y + s.L[1]) < fabs(y) ? y + s.L[1] : y;
        y = fabs(y - s.L[1]) < fabs(y) ? y - s.L[1] : y;
    }
    if (s.per[2]) {
        z = fabs(z + s.L[2]) < fabs(z) ? z + s.L[2] : z;
        z = fabs(z - s.L[2]) < fabs(z) ? z - s.L[2] : z;
    }
}

// |d|^2 summed left to right with no contraction, as the reference's
// scalar code does (verlet-list.cxx:399, LJs-cut.cxx:46).
__device__ inline double sq3_exact(double x, double y, double z) {
#pragma clang fp contract(off)
    return x * x + y * y + z * z;
}

template <typename T>
__device__ inline T wave_sum_group(T v, int tpa) {
    for (int o = tpa >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Pair image rule of the reference's Periodic boundary
// (solver/boundary/periodic.cxx:36-60 + the direct pass of pair.cxx:72-86):
// the raw difference is used when |raw|^2 < max(rns)^2, otherwise the
// single-shift minimum image, dropped when that is beyond max(rns).
// Returns false when the pair has no image.
__device__ inline bool pair_image(const SysDesc& s, double& dx, double& dy, double& dz, double& d2) {
    d2 = sq3_exact(dx, dy, dz);
    if (!s.any_per || d2 < s.rns2) return true;
    shortest3(s, dx, dy, dz);
    d2 = sq3_exact(dx, dy, dz);
    return !(d2 > s.rns2);
}

struct AtomSlots {
    int i;    // atom index (sorted order)
    int sub;  // lane within the atom's group
    long long base;  // list index of chunk 0 for this lane
};

template <int TPA>
__device__ inline AtomSlots atom_slots(int M) {
    const int lane = threadIdx.x & (kWave - 1);
    const long long gw = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    AtomSlots a;
    a.i = (int)(gw * (kWave / TPA) + lane / TPA);
    a.sub = lane % TPA;
    a.base = gw * (long long)M * kWave + lane;
    return a;
}

// list index of slot k of atom i for lists packed with TPA lanes per atom
__host__ __device__ inline long long list_index(int i, int k, int tpa, int M) {
    const int A = kWave / tpa;
    const long long w = i / A;
    const int l = (i % A) * tpa + (k % tpa);
    return (w * M + k / tpa) * kWave + l;
}

}  // namespace mdc
