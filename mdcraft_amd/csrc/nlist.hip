// Neighbour maintenance on the device: wrap + linked-cell binning as an
// atomic-count / prefix-scan counting sort, a deterministic within-cell
// order, a physical reorder of every per-atom array into cell order, and the
// Verlet-list build in the wavefront-packed layout.
//
// Replaces neibs::VerletList::update / build_linked_cells / are_neibs and
// decide (mdcraft/neibs/verlet-list.cxx:41-55, :110-158, :306-400) and the
// grid of mdcraft/neibs/grid.cxx.  The pair set is the reference's: all j
// (self excluded here, added back on export) with
// |d|^2 <= max(rns_i, rns_j)^2 where d = r_i - r_j gets one shift of the box
// length on a periodic axis when 2|d| exceeds it, summed in FP64 without
// contraction.  The reference applies that shift only for atoms in edge
// cells; for three or more cells a shift from an interior cell can never
// produce a pair and for one or two cells every cell is an edge, so the sets
// are the same.
#include "engine.h"

namespace mdc {

// The per-atom kernels of a rebuild take atoms by XCD the way the step
// kernels do (xcd_block): each XCD binds, reorders and lists the slab of
// cells whose positions its L2 already holds from the last step.

// Runs of equal keys among the active lanes of a wavefront (inactive lanes
// above the last active one, as at the end of an array): whether this lane
// starts a run, the run's length, and the lane of its head.
struct RunOf {
    bool head;
    int len, lead;
};
__device__ inline RunOf run_of(int key, bool valid) {
    const int lane = threadIdx.x & (kWave - 1);
    const int prev = __shfl_up(key, 1, kWave);
    const bool head = valid && (lane == 0 || prev != key);
    const unsigned long long heads = __ballot(head);
    const unsigned long long live = __ballot(valid);
    const unsigned long long upto = (2ull << lane) - 1ull;  // lanes 0..lane (all lanes at 63)
    RunOf r;
    r.head = head;
    r.lead = 63 - __clzll(heads & upto);
    const unsigned long long later = heads & ~upto;
    const int end = later ? __ffsll((long long)later) - 1 : 64 - __clzll(live);
    r.len = end - lane;
    return r;
}

// Exclusive scan of the per-cell counts by one workgroup of 1024 threads;
// also clears the counts and the fill counters for the next rebuild.  start
// has ncells + 1 entries.  (A scan by the binning kernel's last workgroup
// measured slower than this launch.)
__global__ void __launch_bounds__(1024) k_scan_cells(int* __restrict__ count, int* __restrict__ start,
                                                     int* __restrict__ fill, int ncells) {
    __shared__ int warp_tot[16];
    __shared__ int carry;
    const int nt = blockDim.x, nw = nt >> 6;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int base = 0; base < ncells; base += nt) {
        const int idx = base + threadIdx.x;
        const int v = idx < ncells ? count[idx] : 0;
        if (idx < ncells) {
            fill[idx] = 0;
            count[idx] = 0;
        }
        int x = v;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) warp_tot[wv] = x;
        __syncthreads();
        if (wv == 0) {
            int t = lane < nw ? warp_tot[lane] : 0;
            for (int o = 1; o < 16; o <<= 1) {
                const int y = __shfl_up(t, o, 64);
                if (lane >= o) t += y;
            }
            if (lane < nw) warp_tot[lane] = t;
        }
        __syncthreads();
        const int before = carry + (wv > 0 ? warp_tot[wv - 1] : 0) + x - v;
        if (idx < ncells) start[idx] = before;
        __syncthreads();
        if (threadIdx.x == nt - 1) carry = before + v;
        __syncthreads();
    }
    if (threadIdx.x == 0) start[ncells] = carry;
}

// lattice::Domain::fit_in_period (domain.cxx:41-47) and the cell of an atom
// in its system's grid (cells of edge >= max rns spanning the box), counted
// per cell.
__global__ void __launch_bounds__(256) k_wrap_count(double4* __restrict__ pos, const SysDesc* __restrict__ desc,
                                                    int* __restrict__ atom_cell, int* __restrict__ cell_count,
                                                    int* __restrict__ misc, int n) {
    const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i == 0) {
        // a rebuild starts: clear the overflow maximum and the step trigger
        misc[M_MAXNN] = 0;
        misc[M_ABORT] = 0x7fffffff;
    }
    const bool valid = i < n;
    int cid = -1;
    if (valid) {
        double4 p = pos[i];
        const SysDesc& s = desc[w_sys(p.w)];
        if (w_void(p.w)) {
            cid = s.void_cell;
        } else {
            double r[3] = {p.x, p.y, p.z};
            int c[3];
            for (int a = 0; a < 3; ++a) {
                if (s.per[a]) r[a] = r[a] < s.lo[a] ? r[a] + s.L[a] : r[a] > s.hi[a] ? r[a] - s.L[a] : r[a];
                double q = floor((r[a] - s.lo[a]) / s.csize[a]);
                int ci = q < 0.0 ? 0 : (q >= (double)s.nc[a] ? s.nc[a] - 1 : (int)q);
                c[a] = ci;
            }
            p.x = r[0];
            p.y = r[1];
            p.z = r[2];
            pos[i] = p;
            cid = s.cell_off + c[0] + s.nc[0] * (c[1] + s.nc[1] * c[2]);
        }
        atom_cell[i] = cid;
    }
    // atoms come in the previous cell order: one add per run of equal cells
    const RunOf run = run_of(cid, valid);
    if (run.head) atomicAdd(&cell_count[cid], run.len);
}

// Each run of equal cells in a wavefront reserves its slots with one
// atomic, in lane order; the order of runs inside a cell is settled by
// k_permute.
__global__ void k_scatter(const int* __restrict__ atom_cell, const int* __restrict__ start, int* __restrict__ fill,
                          int* __restrict__ src, int n) {
    const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const int c = valid ? atom_cell[i] : -1;
    const RunOf r = run_of(c, valid);
    int base = r.head ? atomicAdd(&fill[c], r.len) : 0;
    base = __shfl(base, r.lead, kWave);
    if (valid) src[start[c] + base + ((int)(threadIdx.x & (kWave - 1)) - r.lead)] = i;
}

// Move the records [0, count) (arrays of stride n) into the new (cell)
// order, ascending previous index inside each cell: the thread of scatter
// slot sp takes atom s = src[sp] to slot start[c] + (the number of the
// cell's atoms with a smaller index), so the scatter's run order never
// shows.  Void records (the void cell) keep their scatter slot: their order
// carries nothing, and ranking a cell of thousands of them would cost the
// square of its size.  Also records the cell of each sorted atom and the
// displacement reference positions, and (uq_out given) the fixed-point
// records of the sorted positions.
__global__ void __launch_bounds__(256) k_permute(const int* __restrict__ src, const int* __restrict__ start,
                                                 const double4* __restrict__ pos_in, double4* __restrict__ pos_out,
                                                 const double* __restrict__ din, double* __restrict__ dout,
                                                 const float* __restrict__ fin, float* __restrict__ fout,
                                                 const int* __restrict__ oin, int* __restrict__ oout,
                                                 const int* __restrict__ cell_in, int* __restrict__ cell_out,
                                                 double4* __restrict__ rsave, int* __restrict__ where, int count, int n,
                                                 int void_cell, uint4* __restrict__ uq_out,
                                                 const SysDesc* __restrict__ desc) {
    const int sp = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (sp >= count) return;
    const int s = src[sp];
    const int c = cell_in[s];
    int p = sp;
    if (c != void_cell) {
        const int b = start[c], e = start[c + 1];
        int rank = 0;
        for (int q = b; q < e; ++q) rank += src[q] < s ? 1 : 0;
        p = b + rank;
    }
    const double4 q = pos_in[s];
    pos_out[p] = q;
    rsave[p] = q;
    // during a fixed-point step sequence: the sorted positions' records
    if (uq_out) uq_out[p] = fixq4(q.x, q.y, q.z, q.w, desc[w_sys(q.w)]);
    for (int f = 0; f < D_NFIELDS; ++f) dout[(size_t)f * n + p] = din[(size_t)f * n + s];
    for (int f = 0; f < F_NFIELDS; ++f) fout[(size_t)f * n + p] = fin[(size_t)f * n + s];
    const int o = oin[s];
    oout[p] = o;
    where[o] = p;
    cell_out[p] = c;
}

// After a sort of all records: the void tail [live, n) of the current
// arrays into the idle ones, so that both position buffers and both field
// copies hold it and sorts of [0, live) and steps leave it as it is.
__global__ void __launch_bounds__(256) k_copy_tail(const double4* __restrict__ pos_a, double4* __restrict__ pos_b,
                                                   const double* __restrict__ da, double* __restrict__ db,
                                                   const float* __restrict__ fa, float* __restrict__ fb,
                                                   const int* __restrict__ oa, int* __restrict__ ob, int live, int n) {
    const long long i = live + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pos_b[i] = pos_a[i];
    for (int f = 0; f < D_NFIELDS; ++f) db[(size_t)f * n + i] = da[(size_t)f * n + i];
    for (int f = 0; f < F_NFIELDS; ++f) fb[(size_t)f * n + i] = fa[(size_t)f * n + i];
    ob[i] = oa[i];
}

// One workgroup per cell.  Wave 0 lays out the 27-cell stencil (with the
// reference's de-duplication for one or two periodic cells,
// grid.cxx:150-166) as concatenated candidate ranges; the workgroup stages
// the candidates' positions and search radii in LDS, kTile at a time; each
// wavefront then takes atoms of the cell and tests them against the staged
// tile 64 candidates at a time with the FP64 rule of
// verlet-list.cxx:374-400, compacting the hits with a ballot so the slots
// keep stencil order, then cell order.  The candidates are gathered once
// per cell instead of once per atom.  Counts beyond the capacity are still
// counted, and an overflow raises the step abort flag for every later step,
// so that no step runs on the short list before the host grows it.
constexpr int kTile = 448;
// Four wavefronts per cell (two measured slower: less latency hidden).
constexpr int kBuildThreads = 256;
constexpr int kBuildWaves = kBuildThreads / kWave;
// atoms a wave tests together on the fast path: two chains hide the FP64
// latency; three or four need more than 64 VGPRs and halve the occupancy
constexpr int kGroup = 2;

// The fast-path test of NA atoms of the cell against one staged tile
// (stencil image bits carried by the candidates): the atoms share each
// candidate's LDS reads and image decode, and their chains are independent.
// SHIFT adds the candidate's image, skipped for a home cell with no stencil
// cell across a periodic face (the shift is then zero and fma(0, L, dx) is
// dx).  ibase[t] is atom t's slot 0 in the wavefront-packed list; lt =
// log2(tpa); cnt[t] is atom t's running count.
template <bool SHIFT, int NA>
__device__ inline void build_tile(const int* c_j, const double* c_x, const double* c_y, const double* c_z,
                                  const double* c_r, int tn, const int* ia, const double4* pa, const double* ra,
                                  double L0, double L1, double L2, int lane, int* cnt, int* __restrict__ nl,
                                  const long long* ibase, int lt, int cap) {
    const int tmask = (1 << lt) - 1;
    for (int q0 = 0; q0 < tn; q0 += kWave) {
        const int q = q0 + lane;
        const bool in = q < tn;
        const int qq = in ? q : 0;
        const unsigned ent = (unsigned)c_j[qq];
        const double cr = c_r[qq], cx = c_x[qq], cy = c_y[qq], cz = c_z[qq];
        const int jj = (int)(ent & kIdxMask);
        double sx = 0.0, sy = 0.0, sz = 0.0;
        if (SHIFT) {
            // 1 - stored bits = the shift of r_i - r_j; (+-1) * L is exact
            sx = (double)(1 - (int)((ent >> 26) & 3u));
            sy = (double)(1 - (int)((ent >> 28) & 3u));
            sz = (double)(1 - (int)((ent >> 30) & 3u));
        }
        bool hit[NA];
#pragma unroll
        for (int t = 0; t < NA; ++t) {
            const double rr = fmax(ra[t], cr);
            const double d2c = rr * rr;
            double dx = pa[t].x - cx, dy = pa[t].y - cy, dz = pa[t].z - cz;
            if (SHIFT) {
                dx = fma(sx, L0, dx);
                dy = fma(sy, L1, dy);
                dz = fma(sz, L2, dz);
            }
            hit[t] = in && jj != ia[t] && sq3_exact(dx, dy, dz) <= d2c;
        }
#pragma unroll
        for (int t = 0; t < NA; ++t) {
            const unsigned long long m = __ballot(hit[t]);
            if (hit[t]) {
                const int at = cnt[t] + __popcll(m & ((1ull << lane) - 1ull));
                if (at < cap) nl[ibase[t] + (long long)(at >> lt) * kWave + (at & tmask)] = (int)ent;
            }
            cnt[t] += __popcll(m);
        }
    }
}

__global__ void __launch_bounds__(kBuildThreads) k_build(const double4* __restrict__ pos, const double* __restrict__ rns,
                                               const int* __restrict__ start, const SysDesc* __restrict__ desc,
                                               int* __restrict__ nl, int* __restrict__ nn, int* __restrict__ misc,
                                               int tpa, int M, int cap, int split) {
    __shared__ int s_beg[32], s_pre[32];
    __shared__ unsigned s_img[32];  // image bits of each stencil cell (fast path)
    __shared__ int s_anyimg;        // some stencil cell lies across a periodic face
    __shared__ int c_j[kTile];
    __shared__ double c_x[kTile], c_y[kTile], c_z[kTile], c_r[kTile];
    __shared__ int s_cnt[64];
    // split workgroups per cell, each with a slice of the cell's atoms (a
    // multiple of the 8 atoms its waves take per pass) and the whole stencil
    const int w = xcd_block(blockIdx.x, gridDim.x);
    const int c = w / split, part = w - c * split;
    const int a0 = start[c], a1 = start[c + 1];
    const int slice = (((a1 - a0) + split - 1) / split + kGroup * kBuildWaves - 1) / (kGroup * kBuildWaves) *
                      (kGroup * kBuildWaves);
    const int b0 = a0 + part * slice, b1 = min(a1, b0 + slice);
    if (b0 >= b1) return;  // uniform over the workgroup
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const SysDesc& s = desc[w_sys(pos[a0].w)];
    if (c == s.void_cell) {
        // void records of empty slot space: no list (uniform over the workgroup)
        for (int a = b0 + tid; a < b1; a += kBuildThreads) nn[a] = 0;
        return;
    }
    if (wv == 0) {
        const int lc = c - s.cell_off;
        const int nx = s.nc[0], ny = s.nc[1], nz = s.nc[2];
        const int cx = lc % nx, cy = (lc / nx) % ny, cz = lc / (nx * ny);
        int cnt_c = 0, beg_c = 0;
        if (lane < 27) {
            const int ox = lane / 9 - 1, oy = (lane / 3) % 3 - 1, oz = lane % 3 - 1;
            bool ok = true;
            int x = cx + ox, y = cy + oy, z = cz + oz;
            // r_i - r_j is near -L across the low face (the reference then
            // adds L: shift +1) and near +L across the high face (shift -1);
            // stored as 1 - shift per axis, the list entry's image bits
            const unsigned ix = s.per[0] ? (x < 0 ? 0u : x >= nx ? 2u : 1u) : 1u;
            const unsigned iy = s.per[1] ? (y < 0 ? 0u : y >= ny ? 2u : 1u) : 1u;
            const unsigned iz = s.per[2] ? (z < 0 ? 0u : z >= nz ? 2u : 1u) : 1u;
            s_img[lane] = (ix << 26) | (iy << 28) | (iz << 30);
            if (s.per[0]) {
                ok = ok && !((nx == 1 && ox != 0) || (nx == 2 && ox == 1));
                x = (x + nx) % nx;
            } else ok = ok && x >= 0 && x < nx;
            if (s.per[1]) {
                ok = ok && !((ny == 1 && oy != 0) || (ny == 2 && oy == 1));
                y = (y + ny) % ny;
            } else ok = ok && y >= 0 && y < ny;
            if (s.per[2]) {
                ok = ok && !((nz == 1 && oz != 0) || (nz == 2 && oz == 1));
                z = (z + nz) % nz;
            } else ok = ok && z >= 0 && z < nz;
            if (ok) {
                const int cc = s.cell_off + x + nx * (y + ny * z);
                beg_c = start[cc];
                cnt_c = start[cc + 1] - beg_c;
            }
        }
        const unsigned long long wrapped = __ballot(lane < 27 && cnt_c > 0 && s_img[lane] != kNoShift);
        if (lane == 0) s_anyimg = wrapped != 0ull;
        int pre = cnt_c;
        for (int o = 1; o < 32; o <<= 1) {
            const int y = __shfl_up(pre, o, kWave);
            if (lane >= o) pre += y;
        }
        if (lane < 32) {
            s_beg[lane] = beg_c;
            s_pre[lane] = pre;  // inclusive; lanes 27..31 repeat the total
        }
    }
    __syncthreads();
    const int total = s_pre[26];
    const double L0 = s.L[0], L1 = s.L[1], L2 = s.L[2];
    const int p0 = s.per[0], p1 = s.per[1], p2 = s.per[2];
    // With at least 5 cells on every periodic axis, the reference's shift
    // test (2 dx < -L, 2 dx > L) holds for every candidate of a wrapped
    // stencil cell and for none of an unwrapped one, with a margin of a
    // whole cell: |2 dx| exceeds L + cell across a face and stays below
    // L - cell inside.  The shift then comes with the candidate, and dx is
    // adjusted by the same exact add of +-L.
    const bool fast = (!p0 || s.nc[0] >= 5) && (!p1 || s.nc[1] >= 5) && (!p2 || s.nc[2] >= 5);
    const bool shifted = s_anyimg;
    int lt = 0;
    while ((1 << lt) < tpa) ++lt;
    const int lA = 6 - lt;  // log2 of the atoms per wavefront
    for (int g0 = b0; g0 < b1; g0 += 64) {
        const int ga = min(64, b1 - g0);
        if (tid < 64) s_cnt[tid] = 0;
        for (int t0 = 0; t0 < total; t0 += kTile) {
            const int tn = min(kTile, total - t0);
            __syncthreads();  // the previous tile is consumed and s_cnt is set
            for (int q = tid; q < tn; q += kBuildThreads) {
                const int idx = t0 + q;
                // stencil cell of candidate idx: the count of inclusive
                // prefixes <= idx (s_pre[27..31] hold the total, > idx)
                int cc = 0;
                for (int step = 16; step > 0; step >>= 1)
                    if (s_pre[cc + step - 1] <= idx) cc += step;
                const int j = s_beg[cc] + idx - (cc > 0 ? s_pre[cc - 1] : 0);
                const double4 pj = pos[j];
                c_j[q] = fast ? (int)((unsigned)j | s_img[cc]) : j;
                c_x[q] = pj.x;
                c_y[q] = pj.y;
                c_z[q] = pj.z;
                c_r[q] = rns[j];
            }
            __syncthreads();
            // the atom's own cell is stencil cell 13, staged from s_pre[12]
            // on: read it from the tile when it is there
            auto atom_in = [&](int a, int& i, double4& pi, double& ri) {
                i = g0 + a;
                const int qi = s_pre[12] + (i - a0) - t0;
                const bool own = qi >= 0 && qi < tn;
                pi = own ? make_double4(c_x[qi], c_y[qi], c_z[qi], 0.0) : pos[i];
                ri = own ? c_r[qi] : rns[i];
            };
            // fast path: kGroup atoms a wave at a time
            for (int a = wv; fast && a < ga; a += kGroup * kBuildWaves) {
                int ia[kGroup], cnt[kGroup];
                double4 pa[kGroup];
                double ra[kGroup];
                long long ibase[kGroup];
                int na = 0;
                for (int t = 0; t < kGroup; ++t) {
                    const int at = a + t * kBuildWaves;
                    ia[t] = 0;
                    pa[t] = make_double4(0, 0, 0, 0);
                    ra[t] = 0.0;
                    cnt[t] = 0;
                    if (at < ga) {
                        atom_in(at, ia[t], pa[t], ra[t]);
                        cnt[t] = s_cnt[at];
                        na = t + 1;
                    }
                    ibase[t] = ((long long)(ia[t] >> lA) * M) * kWave + ((ia[t] & ((1 << lA) - 1)) << lt);
                }
                auto run = [&](auto shift_tag) {
                    constexpr bool SH = decltype(shift_tag)::value;
                    static_assert(kGroup == 2, "one instantiation per group size");
                    if (na == 2)
                        build_tile<SH, 2>(c_j, c_x, c_y, c_z, c_r, tn, ia, pa, ra, L0, L1, L2, lane, cnt, nl, ibase, lt,
                                          cap);
                    else
                        build_tile<SH, 1>(c_j, c_x, c_y, c_z, c_r, tn, ia, pa, ra, L0, L1, L2, lane, cnt, nl, ibase, lt,
                                          cap);
                };
                if (shifted) run(std::true_type{});
                else run(std::false_type{});
                if (lane == 0)
                    for (int t = 0; t < na; ++t) s_cnt[a + t * kBuildWaves] = cnt[t];
            }
            for (int a = wv; !fast && a < ga; a += kBuildWaves) {
                int i;
                double4 pi;
                double ri;
                atom_in(a, i, pi, ri);
                int cnt = s_cnt[a];
                for (int q0 = 0; q0 < tn; q0 += kWave) {
                    const int q = q0 + lane;
                    bool hit = false;
                    unsigned ent = 0;
                    if (q < tn) {
                        const int j = c_j[q];
                        if (j != i) {  // j < 2^26 (checked on the host)
                            const double rr = fmax(ri, c_r[q]);
                            const double d2c = rr * rr;
                            double dx = pi.x - c_x[q], dy = pi.y - c_y[q], dz = pi.z - c_z[q];
                            int sx = 0, sy = 0, sz = 0;
                            if (p0 && (2 * dx < -L0)) dx += L0, sx = 1;
                            if (p0 && (2 * dx > L0)) dx -= L0, sx = -1;
                            if (p1 && (2 * dy < -L1)) dy += L1, sy = 1;
                            if (p1 && (2 * dy > L1)) dy -= L1, sy = -1;
                            if (p2 && (2 * dz < -L2)) dz += L2, sz = 1;
                            if (p2 && (2 * dz > L2)) dz -= L2, sz = -1;
                            hit = sq3_exact(dx, dy, dz) <= d2c;
                            // the image of r_j - r_i is the negated shift of r_i - r_j
                            ent = (unsigned)j | ((unsigned)(1 - sx) << 26) | ((unsigned)(1 - sy) << 28) |
                                  ((unsigned)(1 - sz) << 30);
                        }
                    }
                    const unsigned long long m = __ballot(hit);
                    if (hit) {
                        const int at = cnt + __popcll(m & ((1ull << lane) - 1ull));
                        if (at < cap) nl[list_index(i, at, tpa, M)] = (int)ent;
                    }
                    cnt += __popcll(m);
                }
                if (lane == 0) s_cnt[a] = cnt;
            }
        }
        __syncthreads();
        if (tid < ga) {
            const int cnt = s_cnt[tid];
            nn[g0 + tid] = cnt;
            // the host needs the maximum only when the list overflowed
            if (cnt > cap) {
                atomicMax(&misc[M_MAXNN], cnt);
                atomicMin(&misc[M_ABORT], kAbortAll);
            }
        }
    }
}

// verlet-list.cxx:110-158 decide(): |shortest(r - r_saved)|^2 > (skin/2)^2
__global__ void k_decide(const double4* __restrict__ pos, const double4* __restrict__ rsave,
                         const SysDesc* __restrict__ desc, int* __restrict__ misc, int n) {
    // n: the live records
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 p = pos[i], q = rsave[i];
    const SysDesc& s = desc[w_sys(p.w)];
    double dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    shortest3(s, dx, dy, dz);
    wave_atomic_or(sq3_exact(dx, dy, dz) > s.trig2, &misc[M_DECIDE], 1);
}

static inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

// A list of capacity for maxnn neighbours per atom (a quarter more, in
// whole groups of lanes).
static void grow_list(Batch& b, int maxnn) {
    const int cap = ((maxnn + maxnn / 4 + 8 + b.tpa - 1) / b.tpa) * b.tpa;
    const int A = kWave / b.tpa;
    const size_t waves = (size_t)(b.n + A - 1) / A;
    const size_t elems = waves * (size_t)(cap / b.tpa) * kWave;
    if (b.nl) HIPCHECK(hipFree(b.nl));
    b.nl = nullptr;
    if (hipMalloc(&b.nl, elems * sizeof(int)) != hipSuccess)
        throw Error(E_NOMEM, "out of device memory growing the neighbour list");
    b.cap = cap;
    b.M = cap / b.tpa;
    b.nl_elems = elems;
    b.alloc_gen++;
}

// One workgroup per cell, or, when the cells are too few to fill the
// device (six resident build workgroups per CU), up to one per 8 of a
// cell's atoms on average: each stages the whole stencil and lists its
// slice of the cell's atoms, which shortens the longest workgroup's chain.
static int build_split(const Batch& b) {
    const long long room = 6LL * std::max(b.ctx->cu_count, 1);
    const int by_room = (int)std::min<long long>(room / std::max(b.ncells, 1), 8);
    const int by_atoms = (b.live + 8LL * b.ncells - 1) / (8LL * b.ncells);
    return std::max(1, std::min(by_room, by_atoms));
}

static void launch_build(Batch& b) {
    const int split = build_split(b);
    hipLaunchKernelGGL(k_build, dim3((unsigned)((long long)b.ncells * split)), dim3(kBuildThreads), 0, b.ctx->stream,
                       b.pos[b.cur], b.d(b.sc, D_RNS), b.cell_start, b.d_desc, b.nl, b.nn, b.misc, b.tpa, b.M, b.cap,
                       split);
}

// Build the list again after an overflow: grow it until every atom fits,
// then clear the abort flag the overflow raised.
void nlist_regrow(Batch& b, int maxnn) {
    hipStream_t st = b.ctx->stream;
    while (maxnn > b.cap) {
        grow_list(b, maxnn);
        HIPCHECK(hipMemsetAsync(b.misc + M_MAXNN, 0, sizeof(int), st));
        launch_build(b);
        HIPCHECK(hipMemcpyAsync(b.ctx->h_pin, b.misc, sizeof(int) * M_COUNT, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        maxnn = b.ctx->h_pin[M_MAXNN];
    }
    const int abort_init = 0x7fffffff;
    HIPCHECK(hipMemcpyAsync(b.misc + M_ABORT, &abort_init, sizeof(int), hipMemcpyHostToDevice, st));
}

// sync: wait for the build and grow the list now if it overflowed.
// Otherwise the overflow shows in the counters of the next steps, whose
// kernels the raised abort flag stops (the stepping loop regrows).
void nlist_rebuild(Batch& b, bool sync) {
    hipStream_t st = b.ctx->stream;
    const int n = b.n;
    // all records when some became void or live since the last sort, else
    // the live prefix only (the void tail stays where it is)
    const int count = b.full_sort ? n : b.live;
    if (count == 0) {
        b.list_valid = true;
        b.shifts_stale = false;
        return;
    }
    hipLaunchKernelGGL(k_wrap_count, dim3(nblk(count, 256)), dim3(256), 0, st, b.pos[b.cur], b.d_desc, b.atom_cell,
                       b.cell_count, b.misc, count);
    hipLaunchKernelGGL(k_scan_cells, dim3(1), dim3(1024), 0, st, b.cell_count, b.cell_start, b.cell_fill, b.ncells);
    hipLaunchKernelGGL(k_scatter, dim3(nblk(count, 256)), dim3(256), 0, st, b.atom_cell, b.cell_start, b.cell_fill,
                       b.sorted_src, count);
    const int s0 = b.sc, s1 = b.sc ^ 1;
    // cell_fill is free after the scatter: reuse it for the sorted cell ids
    hipLaunchKernelGGL(k_permute, dim3(nblk(count, 256)), dim3(256), 0, st, b.sorted_src, b.cell_start, b.pos[b.cur],
                       b.pos[b.cur ^ 1], b.dst[s0], b.dst[s1], b.fst[s0], b.fst[s1], b.orig[s0], b.orig[s1],
                       b.atom_cell, b.cell_fill, b.rsave, b.where, count, n, b.ncells - 1,
                       b.qfast ? b.uq[b.cur ^ 1] : nullptr, b.d_desc);
    b.cur ^= 1;
    b.sc = s1;
    if (b.full_sort) {
        int live = 0;
        for (int s = 0; s < b.nsys; ++s) live += b.present[s];
        b.live = live;
        if (live < n)
            hipLaunchKernelGGL(k_copy_tail, dim3(nblk(n - live, 256)), dim3(256), 0, st, b.pos[b.cur],
                               b.pos[b.cur ^ 1], b.dst[s1], b.dst[s0], b.fst[s1], b.fst[s0], b.orig[s1], b.orig[s0],
                               live, n);
        b.full_sort = false;
    }
    launch_build(b);
    if (sync) {
        HIPCHECK(hipMemcpyAsync(b.ctx->h_pin, b.misc, sizeof(int) * M_COUNT, hipMemcpyDeviceToHost, st));
        stream_wait(*b.ctx);
        const int maxnn = b.ctx->h_pin[M_MAXNN];  // zero unless some atom overflowed
        if (maxnn > b.cap) nlist_regrow(b, maxnn);
    }
    b.list_valid = true;
    b.shifts_stale = false;
    b.rebuilds++;
    b.since_rebuild = 0;
}

int nlist_decide(Batch& b, double rcut) {
    hipStream_t st = b.ctx->stream;
    if (!b.list_valid) return 1;
    set_trigger(b, rcut, false);
    HIPCHECK(hipMemsetAsync(b.misc + M_DECIDE, 0, sizeof(int), st));
    if (b.live > 0)
        hipLaunchKernelGGL(k_decide, dim3(nblk(b.live, 256)), dim3(256), 0, st, b.pos[b.cur], b.rsave, b.d_desc,
                           b.misc, b.live);
    HIPCHECK(hipMemcpyAsync(b.ctx->h_pin, b.misc, sizeof(int) * M_COUNT, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    return b.ctx->h_pin[M_DECIDE] != 0;
}

}  // namespace mdc
