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
    // 1 when every periodic edge exceeds 2 (rns + skin): the minimum image
    // of a listed pair is then unique and fixed between rebuilds, and the
    // image shift stored in the list entry gives the reference's image rule
    int shift_ok;
    // the box's last cell (after its nc[0] nc[1] nc[2] grid): void records
    // of the box are binned there and no list build visits it
    int void_cell;
    // 32-bit fixed-point coordinates (set with shift_ok): u = (x - lo)
    // qs mod 2^32 with qs = 2^32 / modulus, the modulus being the edge of
    // a periodic axis (the wrap of the integer difference is then the
    // minimum image) and a power of two above 4 (rns + skin) on an open
    // axis; qf = modulus / 2^32 turns a difference back into nm
    double qs[3];
    float qf[3];
    int qpad;
};

// A list entry holds the neighbour index in bits 0-25 and, per axis, the
// image shift (-1, 0, +1) + 1 of r_j - r_i in bits 26-27, 28-29, 30-31.
constexpr unsigned kIdxMask = (1u << 26) - 1u;
// image bits of an unshifted pair (1 - shift = 1 on every axis)
constexpr unsigned kNoShift = (1u << 26) | (1u << 28) | (1u << 30);
// step abort flag value that stops every step (a list that overflowed)
constexpr int kAbortAll = -0x7fffffff - 1;
constexpr int kMaxAtomsShift = 1 << 26;
__host__ __device__ inline int ent_j(int e) { return (int)((unsigned)e & kIdxMask); }
__host__ __device__ inline int ent_s(int e, int axis) { return (int)(((unsigned)e >> (26 + 2 * axis)) & 3u) - 1; }

// raw buffer loads: 32-bit offsets, out-of-range reads return zero
// without touching memory; kSkip is beyond every buffer here, for lanes
// whose value is not needed
constexpr unsigned kSkip = 0x80000000u;
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
typedef double d2_t __attribute__((ext_vector_type(2)));
__device__ inline __amdgpu_buffer_rsrc_t make_buf(const void* p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ inline double4 buf_d4(__amdgpu_buffer_rsrc_t r, unsigned off) {
    const d2_t a = __builtin_bit_cast(d2_t, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
    const d2_t b = __builtin_bit_cast(d2_t, __builtin_amdgcn_raw_buffer_load_b128(r, off + 16, 0, 0));
    return make_double4(a[0], a[1], b[0], b[1]);
}
// The whole 128-bit vector is reinterpreted at once: per-element casts of
// the loaded vector have been narrowed to a single dword by the compiler.
typedef float f32x4_t __attribute__((ext_vector_type(4)));
__device__ inline float4 buf_f4(__amdgpu_buffer_rsrc_t r, unsigned off) {
    const f32x4_t a = __builtin_bit_cast(f32x4_t, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
    return make_float4(a[0], a[1], a[2], a[3]);
}
__device__ inline double buf_d1(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

// A dword of the adjacent lane (lane 2k <-> 2k + 1), by a DPP quad
// permutation [1, 0, 3, 2]; every lane of the wave must be active.
__device__ inline unsigned swap_adj(unsigned v) {
    return (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0xB1, 0xF, 0xF, false);
}

// Cooperative 32-byte gathers.  Lanes 2k and 2k + 1 each need one 32-byte
// record, at byte offsets off (kSkip: none, read as zeros).  In the first
// load both lanes read the even lane's record, the even lane its low half
// and the odd lane its high half; in the second both read the odd lane's
// record; one exchange of 16 bytes then gives each lane its whole record.
// Both halves of a record come from one instruction, so a record costs one
// cache-line lookup in the vector L1 instead of two, which is what bounds
// neighbour gathers here.  Every lane of the wave must execute it (uniform
// control flow).
__device__ inline void load32_pair(__amdgpu_buffer_rsrc_t r, unsigned off, u32x4_t& lo, u32x4_t& hi) {
    const bool odd = (__lane_id() & 1u) != 0;
    const unsigned other = swap_adj(off);
    const unsigned half = odd ? 16u : 0u;
    const u32x4_t a = __builtin_amdgcn_raw_buffer_load_b128(r, (odd ? other : off) + half, 0, 0);
    const u32x4_t b = __builtin_amdgcn_raw_buffer_load_b128(r, (odd ? off : other) + half, 0, 0);
    const u32x4_t send = odd ? a : b;
    u32x4_t recv;
    recv[0] = swap_adj(send[0]);
    recv[1] = swap_adj(send[1]);
    recv[2] = swap_adj(send[2]);
    recv[3] = swap_adj(send[3]);
    lo = odd ? recv : a;
    hi = odd ? b : recv;
}
__device__ inline double4 buf_d4_pair(__amdgpu_buffer_rsrc_t r, unsigned off) {
    u32x4_t lo, hi;
    load32_pair(r, off, lo, hi);
    const d2_t a = __builtin_bit_cast(d2_t, lo), b = __builtin_bit_cast(d2_t, hi);
    return make_double4(a[0], a[1], b[0], b[1]);
}
// two float4 records side by side (32 bytes), read the same way
__device__ inline void buf_f4x2_pair(__amdgpu_buffer_rsrc_t r, unsigned off, float4& first, float4& second) {
    u32x4_t lo, hi;
    load32_pair(r, off, lo, hi);
    const f32x4_t a = __builtin_bit_cast(f32x4_t, lo), b = __builtin_bit_cast(f32x4_t, hi);
    first = make_float4(a[0], a[1], a[2], a[3]);
    second = make_float4(b[0], b[1], b[2], b[3]);
}

// tag word packed into pos.w: kind (bits 0-7), skipped (8), void (9: a
// record of a box slot that holds no atom; it is binned into the box's void
// cell, which no list build visits, and is skipped as a centre), system
// (bits 10 and up)
__host__ __device__ inline double pack_w(int kind, int inactive, int sys, int voided = 0) {
    return (double)((long long)kind | ((long long)inactive << 8) | ((long long)voided << 9) | ((long long)sys << 10));
}
__device__ inline int w_kind(double w) { return (int)((long long)w & 0xff); }
__device__ inline int w_inactive(double w) { return (int)(((long long)w >> 8) & 1); }
__device__ inline int w_void(double w) { return (int)(((long long)w >> 9) & 1); }
__device__ inline int w_sys(double w) { return (int)((long long)w >> 10); }

// The part of a system's descriptor that the pair loops read, copied into
// registers once per atom so it is not reloaded per neighbour.
struct Frame {
    double L[3];
    double rns2;
    int per[3];
    int any_per;
    int shift_ok;
};
__device__ inline Frame frame_of(const SysDesc& s) {
    Frame f;
    for (int c = 0; c < 3; ++c) {
        f.L[c] = s.L[c];
        f.per[c] = s.per[c];
    }
    f.rns2 = s.rns2;
    f.any_per = s.any_per;
    f.shift_ok = s.shift_ok;
    return f;
}

// lattice::Domain::shortest (mdcraft/lattice/domain.cxx:59-66)
template <class S>
__device__ inline void shortest3(const S& s, double& x, double& y, double& z) {
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

// Fixed-point record of a position: three 32-bit coordinates (SysDesc qs)
// and the low 32 bits of the tag word.  A listed pair's difference is then
// (int)(u_j - u_i) * qf per axis: the integer wrap takes the minimum image
// on periodic axes, exact to one unit (about 1e-8 of the edge).
__device__ inline unsigned fixq1(double x, double lo, double qs) {
    return (unsigned)(long long)rint((x - lo) * qs);
}
__device__ inline uint4 fixq4(double x, double y, double z, double w, const SysDesc& s) {
    return make_uint4(fixq1(x, s.lo[0], s.qs[0]), fixq1(y, s.lo[1], s.qs[1]), fixq1(z, s.lo[2], s.qs[2]),
                      (unsigned)(long long)w);
}
__device__ inline float fixq_d(unsigned uj, unsigned ui, float qf) { return (float)(int)(uj - ui) * qf; }
__device__ inline uint4 buf_u4(__amdgpu_buffer_rsrc_t r, unsigned off) {
    const u32x4_t a = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    return make_uint4(a[0], a[1], a[2], a[3]);
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
template <class S>
__device__ inline bool pair_image(const S& s, double& dx, double& dy, double& dz, double& d2) {
    d2 = sq3_exact(dx, dy, dz);
    if (!s.any_per || d2 < s.rns2) return true;
    shortest3(s, dx, dy, dz);
    d2 = sq3_exact(dx, dy, dz);
    return !(d2 > s.rns2);
}

// Raise a flag word once per wavefront: every lane with p set would store
// the same value, so the lowest such lane does it for all (atomics from many
// lanes to one address serialise).
__device__ inline void wave_atomic_min(bool p, int* addr, int v) {
    const unsigned long long m = __ballot(p);
    if (p && (int)__lane_id() == __builtin_ctzll(m)) atomicMin(addr, v);
}
__device__ inline void wave_atomic_or(bool p, int* addr, int v) {
    const unsigned long long m = __ballot(p);
    if (p && (int)__lane_id() == __builtin_ctzll(m)) atomicOr(addr, v);
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  xcd_block maps the hardware workgroup index to a logical one so that
// each XCD runs one contiguous range of logical workgroups: for kernels over
// cell-ordered atoms that is one spatial slab, whose neighbours then stay in
// that XCD's L2.  The map is a bijection of [0, G).
constexpr int kXcd = 8;
__device__ inline int xcd_block(int b, int G) {
    const int x = b % kXcd, k = b / kXcd;
    const int per = G / kXcd, rem = G % kXcd;
    return x < rem ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
}

struct AtomSlots {
    int i;    // atom index (sorted order)
    int sub;  // lane within the atom's group
    long long base;  // list index of chunk 0 for this lane
};

template <int TPA>
__device__ inline AtomSlots atom_slots(int M) {
    const int lane = threadIdx.x & (kWave - 1);
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const long long gw = ((long long)blk * blockDim.x + threadIdx.x) / kWave;
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
