// C ABI of the engine (include/mdc.h): contexts, potentials, resident
// systems, upload/download of the reference's 208-byte atom records, and the
// entry points that replace VerletList::update/decide, Solver prepare/stages/
// virials and Stepper::make_step at whole-array granularity.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <numeric>
#include <sstream>
#include "abi.h"

using namespace mdc;

constexpr size_t kNoAllocBytes = (size_t)64 << 20;
static __device__ __align__(256) char scratch[kNoAllocBytes];

namespace {

thread_local std::string g_last;

// Work area of mdc_compute_noalloc: one copy per GPU, held until the process
// ends.  In host code `scratch` is only a stand-in for the device symbol;
// its address on each device is looked up once, when the first context on
// that device is created, and only that address is handed to kernels.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

// the device must be current
void scratch_lookup(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error(E_INVAL, "device index above the supported range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(E_HIP, "no device address for the allocation-free work area");
    g_scratch_ptr[device] = static_cast<char*>(p);
}

char* scratch_on(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}

template <class F>
int guard(Ctx* c, F&& f) {
    try {
        f();
        if (c) c->err.clear();
        return 0;
    } catch (const Error& e) {
        g_last = e.what();
        if (c) c->err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last = "host allocation failed";
        if (c) c->err = g_last;
        return E_NOMEM;
    } catch (const std::exception& e) {
        g_last = e.what();
        if (c) c->err = e.what();
        return E_INVAL;
    }
}

template <class T>
T* dalloc(size_t count) {
    if (count == 0) count = 1;
    void* p = nullptr;
    if (hipMalloc(&p, count * sizeof(T)) != hipSuccess)
        throw Error(E_NOMEM, "out of device memory (" + std::to_string(count * sizeof(T)) + " bytes)");
    return static_cast<T*>(p);
}

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

// The batch's device staging area of at least `bytes` (kept between
// transfers; every transfer waits for its copies before it returns).
char* stage_area(Batch& b, size_t bytes) {
    if (b.stage_bytes < bytes) {
        dfree(b.stage);
        b.stage = dalloc<char>(bytes);
        b.stage_bytes = bytes;
    }
    return b.stage;
}

// Knot slopes of the tables' cubic Hermite splines (EAM.cxx:603-686).  With
// h the knot spacing (H1 = 1/h) the rows are
//   2 z[0] + z[1] = 3 (y[1] - y[0]) / h,
//   z[k-1] / 2 + 2 z[k] + z[k+1] / 2 = 3 (y[k+1] - y[k-1]) / (2 h), 0 < k < N-1,
//   end 1: z[N-1] = 0;  end 2: z[N-2] + 2 z[N-1] = 3 (y[N-1] - y[N-2]) / h.
// The reference's sweep leaves z[0 .. N-2] as the solution of the first N-1
// rows with z[N-1] taken as zero, and z[N-1] as the solution of the end row
// with z[N-2] written through it; with end 1 that is the exact solution,
// with end 2 the slopes below the end differ from it, and this does the
// same.  Rows 0 .. N-2 are reduced from the top (z[k] = m[k] z[k+1] + q[k])
// and solved from z[N-1] = 0 upwards.
void spline_slopes(const double* y, double* z, double H1, size_t N, int end) {
    std::vector<double> m(N), q(N);
    auto row = [&](size_t k, double& lo, double& di, double& up, double& r) {
        if (k == 0) {
            lo = 0.0, di = 2.0, up = 1.0, r = 3.0 * H1 * (y[1] - y[0]);
        } else if (k + 1 < N) {
            lo = 0.5, di = 2.0, up = 0.5, r = 1.5 * H1 * (y[k + 1] - y[k - 1]);
        } else if (end == 1) {
            lo = 0.0, di = 2.0, up = 0.0, r = 0.0;
        } else {
            lo = 1.0, di = 2.0, up = 0.0, r = 3.0 * H1 * (y[N - 1] - y[N - 2]);
        }
    };
    double mp = 0.0, qp = 0.0;  // m[k-1], q[k-1]
    for (size_t k = 0; k + 1 < N; ++k) {
        double lo, di, up, r;
        row(k, lo, di, up, r);
        const double piv = di + lo * mp;
        m[k] = -up / piv;
        q[k] = (r - lo * qp) / piv;
        mp = m[k];
        qp = q[k];
    }
    double zn = 0.0;  // z[N-1] as the upper rows see it
    for (size_t k = N - 1; k-- > 0;) {
        z[k] = m[k] * zn + q[k];
        zn = z[k];
    }
    double lo, di, up, r;
    row(N - 1, lo, di, up, r);
    // z[N-2] = m[N-2] z[N-1] + q[N-2] in the end row
    z[N - 1] = N > 1 ? (r - lo * q[N - 2]) / (di + lo * m[N - 2]) : 0.0;
}

// interval records (y0, z0, y1, z1); the last record repeats the last knot
void records(const double* y, const double* z, size_t N, std::vector<double>& out) {
    for (size_t k = 0; k < N; ++k) {
        size_t k1 = k + 1 < N ? k + 1 : k;
        out.push_back(y[k]);
        out.push_back(z[k]);
        out.push_back(y[k1]);
        out.push_back(z[k1]);
    }
}

// One EAM element (nk = 1, EAM.cxx:15-53) or the nk elements of an alloy
// EAM: per element the rho records [nr], per pair of elements (setfl order
// (1,1), (2,1), (2,2), ...) the r*phi records [nr], per element the
// embedding records [nrho].
Pot* make_eam(Ctx* ctx, int nk, double rcut, const double* F, const double* rho, const double* rphi, int nr, double dr,
              int nrho, double drho) {
    if (nk < 1 || nk > kMaxMultiKinds) throw Error(E_INVAL, "alloy elements must be 1..4");
    if (nr < 3 || nrho < 3) throw Error(E_INVAL, "EAM tables need at least 3 knots");
    if (!(dr > 0) || !(drho > 0) || !(rcut > 0)) throw Error(E_INVAL, "EAM steps and cutoff must be positive");
    auto p = std::make_unique<Pot>();
    p->ctx = ctx;
    p->type = PT_EAM;
    p->rcut = rcut;
    EamTables& T = p->eam;
    T.nk = nk;
    T.nr = nr;
    T.nrho = nrho;
    T.dr = dr;
    T.drho = drho;
    T.rc = rcut;
    const int npair = nk * (nk + 1) / 2;
    std::vector<double> rec64, F64, z(std::max(nr, nrho));
    for (int k = 0; k < nk; ++k) {
        spline_slopes(rho + (size_t)k * nr, z.data(), 1.0 / dr, nr, 1);
        records(rho + (size_t)k * nr, z.data(), nr, rec64);
    }
    for (int q = 0; q < npair; ++q) {
        spline_slopes(rphi + (size_t)q * nr, z.data(), 1.0 / dr, nr, 1);
        records(rphi + (size_t)q * nr, z.data(), nr, rec64);
    }
    for (int k = 0; k < nk; ++k) {
        spline_slopes(F + (size_t)k * nrho, z.data(), 1.0 / drho, nrho, 2);
        records(F + (size_t)k * nrho, z.data(), nrho, F64);
    }
    for (double x : rec64)
        if (!std::isfinite(x)) throw Error(E_RANGE, "non-finite value in the EAM r tables");
    // FP32 records: the interval cubic as coefficients in t = r/dr - N1
    std::vector<float> rec32(rec64.size());
    for (size_t q = 0; q < rec64.size(); q += 4) {
        const double y0 = rec64[q], z0 = rec64[q + 1], y1 = rec64[q + 2], z1 = rec64[q + 3];
        const double sl = (y1 - y0) / dr;
        const double aa = (z1 - sl) - (sl - z0);
        const double bb = aa - (sl - z0);
        rec32[q] = (float)y0;
        rec32[q + 1] = (float)(dr * z0);
        rec32[q + 2] = (float)(-dr * bb);
        rec32[q + 3] = (float)(dr * aa);
    }
    const size_t nrec = rec64.size() / 4;
    if (nrec >= (1ull << 27)) throw Error(E_INVAL, "EAM tables too large");
    T.nrec = nrec;
    T.nFrec = F64.size() / 4;
    T.r64 = dalloc<double4>(nrec);
    T.r32 = dalloc<float4>(nrec);
    T.F64 = dalloc<double4>(F64.size() / 4);
    HIPCHECK(hipMemcpy(T.r64, rec64.data(), rec64.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(T.r32, rec32.data(), rec32.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(T.F64, F64.data(), F64.size() * sizeof(double), hipMemcpyHostToDevice));
    if (nk == 1) {
        std::vector<float> pr((size_t)8 * nr);
        for (int k = 0; k < nr; ++k)
            for (int c = 0; c < 4; ++c) {
                pr[(size_t)8 * k + c] = rec32[(size_t)4 * k + c];
                pr[(size_t)8 * k + 4 + c] = rec32[(size_t)4 * (nr + k) + c];
            }
        T.r32p = dalloc<float4>((size_t)2 * nr);
        HIPCHECK(hipMemcpy(T.r32p, pr.data(), pr.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    if (nk > 1) {
        // one object per pair of elements lo <= hi, as the Multi kernels
        // take them: the lo atom sums rho_hi and embeds with F_lo, the hi
        // atom sums rho_lo and embeds with F_hi
        std::vector<EamObj> h(npair);
        for (int hi = 0; hi < nk; ++hi)
            for (int lo = 0; lo <= hi; ++lo) {
                const int q = pair_index(hi, lo);
                EamObj& o = h[q];
                o.rho_off = hi * nr;
                o.rho_hi_off = lo * nr;
                o.phi_off = (nk + q) * nr;
                o.F_off = lo * nrho;
                o.F_hi_off = hi * nrho;
                o.nr = nr;
                o.nrho = nrho;
                o.pair_only = 0;
                o.dr = dr;
                o.dr1 = 1.0 / dr;
                o.drho = drho;
                o.drho1 = 1.0 / drho;
                o.rc2 = rcut * rcut;
            }
        T.npair = npair;
        T.alloy = true;
        T.objs = dalloc<EamObj>(npair);
        HIPCHECK(hipMemcpy(T.objs, h.data(), sizeof(EamObj) * npair, hipMemcpyHostToDevice));
    }
    return p.release();
}

// Single-element setfl reader with the reference's unit conversion
// (EAM.cxx:55-87): skip 4 lines; Nrho drho Nr dr rc; Z M a lattice; F (eV),
// rho, r*phi (eV*Angstrom); r, rc from Angstrom to nm.
Pot* make_eam_file(Ctx* ctx, const char* path) {
    std::ifstream f(path);
    if (!f.is_open()) throw Error(E_FILE, std::string("Failed to open file: ") + path);
    std::string line, lattice;
    for (int i = 0; i < 4; ++i) std::getline(f, line);
    int nrho = 0, nr = 0, Z = 0;
    double drho = 0, dr = 0, rc = 0, M = 0, a0 = 0;
    f >> nrho >> drho >> nr >> dr >> rc;
    f >> Z >> M >> a0 >> lattice;
    if (!f || nrho < 3 || nr < 3) throw Error(E_FILE, std::string("malformed setfl header: ") + path);
    const double Na = 6.02214076e+23, Qe = 1.602176634e-19;
    const double UnMass = 1e-3 / Na, UnLeng = 1e-9, UnTime = 1e-12;
    const double UnEnrg = UnMass * (UnLeng / UnTime) * (UnLeng / UnTime);
    std::vector<double> F(nrho), rho(nr), rphi(nr);
    for (auto& x : F) {
        f >> x;
        x *= Qe / UnEnrg;
    }
    for (auto& x : rho) f >> x;
    for (auto& x : rphi) {
        f >> x;
        x *= Qe / (10 * UnEnrg);
    }
    if (!f) throw Error(E_FILE, std::string("truncated setfl tables: ") + path);
    return make_eam(ctx, 1, rc / 10, F.data(), rho.data(), rphi.data(), nr, dr / 10, nrho, drho);
}

// N-element setfl ("eam/alloy", the format of problems/NiAl/*.eam.alloy):
// 3 comment lines; "Nelements El1 El2 ..."; Nrho drho Nr dr rc; per element
// "Z mass a lattice", F (eV) on the rho grid, rho on the r grid; then r*phi
// (eV*Angstrom) per pair of elements in the order (1,1), (2,1), (2,2), ...
// Units become the reference's as in EAM.cxx:55-87.  One element gives the
// reference's EAM object.
Pot* make_alloy_file(Ctx* ctx, const char* path) {
    std::ifstream f(path);
    if (!f.is_open()) throw Error(E_FILE, std::string("Failed to open file: ") + path);
    std::string line;
    for (int i = 0; i < 3; ++i) std::getline(f, line);
    std::getline(f, line);
    std::istringstream head(line);
    int ne = 0, nrho = 0, nr = 0;
    double drho = 0, dr = 0, rc = 0;
    head >> ne;
    f >> nrho >> drho >> nr >> dr >> rc;
    if (!head || !f || ne < 1 || nrho < 3 || nr < 3) throw Error(E_FILE, std::string("malformed setfl header: ") + path);
    if (ne > kMaxMultiKinds) throw Error(E_INVAL, "alloy elements must be 1..4");
    const double Na = 6.02214076e+23, Qe = 1.602176634e-19;
    const double UnMass = 1e-3 / Na, UnLeng = 1e-9, UnTime = 1e-12;
    const double UnEnrg = UnMass * (UnLeng / UnTime) * (UnLeng / UnTime);
    std::vector<double> F((size_t)ne * nrho), rho((size_t)ne * nr), rphi((size_t)ne * (ne + 1) / 2 * nr);
    for (int e = 0; e < ne; ++e) {
        int Z = 0;
        double M = 0, a0 = 0;
        std::string lattice;
        f >> Z >> M >> a0 >> lattice;
        for (int k = 0; k < nrho; ++k) {
            f >> F[(size_t)e * nrho + k];
            F[(size_t)e * nrho + k] *= Qe / UnEnrg;
        }
        for (int k = 0; k < nr; ++k) f >> rho[(size_t)e * nr + k];
    }
    for (auto& x : rphi) {
        f >> x;
        x *= Qe / (10 * UnEnrg);
    }
    if (!f) throw Error(E_FILE, std::string("truncated setfl tables: ") + path);
    return make_eam(ctx, ne, rc / 10, F.data(), rho.data(), rphi.data(), nr, dr / 10, nrho, drho);
}

void free_pot(Pot* p) {
    if (!p) return;
    if (p->type == PT_BPNN) {
        bpnn_destroy(p);
        return;
    }
    dfree(p->eam.r64);
    dfree(p->eam.r32);
    dfree(p->eam.r32p);
    dfree(p->eam.F64);
    dfree(p->eam.objs);
    delete p;
}

// solver::Multi over EAM objects (multi.cxx:11-151): the records of every
// object are copied into one array so one kernel serves all pairs of kinds.
Pot* make_multi(Ctx* ctx, int nk, int npots, const int* ki, const int* kj, const Pot* const* pots) {
    if (nk < 1 || nk > kMaxMultiKinds) throw Error(E_INVAL, "Multi kinds must be 1..4");
    const int npair = nk * (nk + 1) / 2;
    std::vector<const Pot*> obj(npair, nullptr);
    for (int q = 0; q < npots; ++q) {
        if (ki[q] < 0 || kj[q] < 0 || ki[q] >= nk || kj[q] >= nk) throw Error(E_INVAL, "Multi: kind out of range");
        const Pot* s = pots[q];
        const bool eam1 = s && s->type == PT_EAM && s->eam.nk == 1 && !s->eam.alloy;
        const bool tab = s && s->type == PT_LJ && s->lj.fn == 2;
        if (!eam1 && !tab)
            throw Error(E_INVAL, "Multi: every member must be a single-element EAM or a tabulated Pair potential");
        obj[pair_index(ki[q], kj[q])] = s;  // a later add_potential replaces an earlier one
    }
    for (auto* o : obj)
        if (!o) throw Error(E_INVAL, "Multi: a pair of kinds has no potential");
    auto p = std::make_unique<Pot>();
    p->ctx = ctx;
    p->type = PT_EAM;
    EamTables& T = p->eam;
    T.nk = nk;
    T.npair = npair;
    std::vector<EamObj> h(npair);
    size_t nrec = 0, nF = 0;
    double rc = 0;
    // A tabulated Pair member becomes an object whose density records are
    // zero (rho' = 0: the pair term alone acts) and whose embedding table
    // is three zero knots on a step so wide that every density falls in
    // its first interval (F = F' = 0).
    const int kZeroF = 3;
    for (int q = 0; q < npair; ++q) {
        const EamTables& S = obj[q]->eam;
        const bool tab = obj[q]->type == PT_LJ;
        EamObj& o = h[q];
        o.rho_off = (int)nrec;
        o.phi_off = (int)(nrec + S.nr);
        o.F_off = (int)nF;
        o.rho_hi_off = o.rho_off;
        o.F_hi_off = o.F_off;
        o.nr = S.nr;
        o.nrho = tab ? kZeroF : S.nrho;
        o.pair_only = tab ? 1 : 0;
        o.dr = S.dr;
        o.dr1 = 1.0 / S.dr;
        o.drho = tab ? 1e300 : S.drho;
        o.drho1 = 1.0 / o.drho;
        o.rc2 = S.rc * S.rc;
        nrec += 2 * (size_t)S.nr;
        nF += (size_t)o.nrho;
        rc = std::max(rc, S.rc);
    }
    if (nrec >= (1ull << 27)) throw Error(E_INVAL, "Multi: tables too large");
    T.r64 = dalloc<double4>(nrec);
    T.r32 = dalloc<float4>(nrec);
    T.F64 = dalloc<double4>(nF);
    T.objs = dalloc<EamObj>(npair);
    for (int q = 0; q < npair; ++q) {
        const EamTables& S = obj[q]->eam;
        if (h[q].pair_only) {
            HIPCHECK(hipMemset(T.r64 + h[q].rho_off, 0, sizeof(double4) * S.nr));
            HIPCHECK(hipMemset(T.r32 + h[q].rho_off, 0, sizeof(float4) * S.nr));
            HIPCHECK(hipMemcpy(T.r64 + h[q].phi_off, S.r64, sizeof(double4) * S.nr, hipMemcpyDeviceToDevice));
            HIPCHECK(hipMemcpy(T.r32 + h[q].phi_off, S.r32, sizeof(float4) * S.nr, hipMemcpyDeviceToDevice));
            HIPCHECK(hipMemset(T.F64 + h[q].F_off, 0, sizeof(double4) * kZeroF));
            continue;
        }
        HIPCHECK(hipMemcpy(T.r64 + h[q].rho_off, S.r64, sizeof(double4) * 2 * S.nr, hipMemcpyDeviceToDevice));
        HIPCHECK(hipMemcpy(T.r32 + h[q].rho_off, S.r32, sizeof(float4) * 2 * S.nr, hipMemcpyDeviceToDevice));
        HIPCHECK(hipMemcpy(T.F64 + h[q].F_off, S.F64, sizeof(double4) * S.nrho, hipMemcpyDeviceToDevice));
    }
    HIPCHECK(hipMemcpy(T.objs, h.data(), sizeof(EamObj) * npair, hipMemcpyHostToDevice));
    T.nrec = nrec;
    T.nFrec = nF;
    // the single-object fields mirror object 0 (unused by the Multi kernels)
    T.nr = h[0].nr;
    T.nrho = h[0].nrho;
    T.dr = h[0].dr;
    T.drho = h[0].drho;
    T.rc = rc;
    p->rcut = rc;
    return p.release();
}

// ---- systems -------------------------------------------------------------

__global__ void k_scatter_in(const double* __restrict__ stage, const double4* __restrict__ spos,
                             const int* __restrict__ where, double4* __restrict__ pos, double* __restrict__ d,
                             float* __restrict__ fl, uint32_t dmask, int set_r, int set_f32, int n) {
    int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n) return;
    int p = where[o];
    double4 q = pos[p];
    double4 s = spos[o];
    if (set_r) {
        q.x = s.x;
        q.y = s.y;
        q.z = s.z;
    }
    q.w = s.w;
    pos[p] = q;
    for (int f = 0; f < D_NFIELDS; ++f)
        if (dmask & (1u << f)) d[(size_t)f * n + p] = stage[(size_t)f * n + o];
    if (set_f32)
        for (int c = 0; c < 3; ++c) fl[(size_t)c * n + p] = (float)stage[(size_t)(D_FX + c) * n + o];
}

// The records of one box, [first, first + cap) of the caller's order, each
// at its sorted place where[first + q]: the first `count` from a staging
// area of stride count (every field, the position and its tag word), the
// rest void (position void_pos with its tag word, every field zero).
__global__ void k_scatter_range(const double* __restrict__ stage, const double4* __restrict__ spos,
                                const int* __restrict__ where, double4* __restrict__ pos, double* __restrict__ d,
                                float* __restrict__ fl, int first, int count, int cap, int n, double4 void_pos) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= cap) return;
    const int p = where[first + q];
    if (q < count) {
        pos[p] = spos[q];
        for (int f = 0; f < D_NFIELDS; ++f) d[(size_t)f * n + p] = stage[(size_t)f * count + q];
        for (int c = 0; c < 3; ++c) fl[(size_t)c * n + p] = (float)stage[(size_t)(D_FX + c) * count + q];
    } else {
        pos[p] = void_pos;
        for (int f = 0; f < D_NFIELDS; ++f) d[(size_t)f * n + p] = 0.0;
        for (int c = 0; c < F_NFIELDS; ++c) fl[(size_t)c * n + p] = 0.0f;
    }
}

// records [first, first + count) of the caller's order into a staging area
__global__ void k_gather_out(double* __restrict__ stage, double4* __restrict__ spos, const int* __restrict__ where,
                             const double4* __restrict__ pos, const double* __restrict__ d, const float* __restrict__ fl,
                             int f_mixed, int n, int first, int count) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    int p = where[first + q];
    spos[q] = pos[p];
    for (int f = 0; f < D_NFIELDS; ++f) stage[(size_t)f * count + q] = d[(size_t)f * n + p];
    if (f_mixed)
        for (int c = 0; c < 3; ++c) stage[(size_t)(D_FX + c) * count + q] = (double)fl[(size_t)c * n + p];
}

// set or clear the skip flag of the atoms [first, first + count) of the
// caller's order in the current positions
__global__ void k_set_inactive(double4* __restrict__ pos, const int* __restrict__ where, int first, int count,
                               int inactive) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const int p = where[first + q];
    double4 v = pos[p];
    v.w = pack_w(w_kind(v.w), inactive, w_sys(v.w), w_void(v.w));
    pos[p] = v;
}

// New positions: the least pair distance the fixed-point passes met is
// forgotten.
static void reset_rmin(Batch& b) {
    static const int none = kRminNone;
    HIPCHECK(hipMemcpyAsync(b.misc + M_RMIN, &none, sizeof(int), hipMemcpyHostToDevice, b.ctx->stream));
}

void free_batch(Batch& b) {
    dfree(b.pos[0]);
    dfree(b.pos[1]);
    dfree(b.dst[0]);
    dfree(b.dst[1]);
    dfree(b.fst[0]);
    dfree(b.fst[1]);
    dfree(b.orig[0]);
    dfree(b.orig[1]);
    dfree(b.rsave);
    dfree(b.posF);
    dfree(b.uq[0]);
    dfree(b.uq[1]);
    dfree(b.uqF);
    b.uq_n = 0;
    dfree(b.where);
    dfree(b.cell_count);
    dfree(b.cell_start);
    dfree(b.cell_fill);
    dfree(b.atom_cell);
    dfree(b.sorted_src);
    dfree(b.nl);
    dfree(b.nn);
    dfree(b.misc);
    dfree(b.fpm);
    dfree(b.d_desc);
    dfree(b.thermo);
    dfree(b.stage);
    b.stage_bytes = 0;
    graph_release(b);
    bpnn_release_work(b);
    for (auto& t : b.ev_pending) {
        b.ev_free.push_back(t.a);
        b.ev_free.push_back(t.b);
    }
    for (auto e : b.ev_free) (void)hipEventDestroy(e);
    b.ev_free.clear();
    b.ev_pending.clear();
}

// Cell grid of each system: cells of edge >= max rns over the box (capped
// per axis), cell offsets into the global cell arrays, and one void cell
// after every box's grid for the void records of all boxes (which the sort
// therefore puts after every live record).
void layout_cells(Batch& b) {
    int off = 0;
    for (int s = 0; s < b.nsys; ++s) {
        SysDesc& d = b.desc[s];
        const double rmax = b.rns_max[s];
        for (int a = 0; a < 3; ++a) {
            int n = 1;
            if (rmax > 0 && d.L[a] > 0) {
                double q = std::floor(d.L[a] / (rmax * (1.0 + 1e-9)));
                n = (int)std::max(1.0, std::min(q, 256.0));
            }
            d.nc[a] = n;
            d.csize[a] = d.L[a] > 0 ? d.L[a] / n : 1e300;
        }
        d.rns2 = rmax * rmax;
        d.cell_off = off;
        off += d.nc[0] * d.nc[1] * d.nc[2];
    }
    for (int s = 0; s < b.nsys; ++s) b.desc[s].void_cell = off;
    off += 1;
    if (off > b.cell_alloc) {
        // grown only: boxes refilled with new geometries change the count
        dfree(b.cell_count);
        dfree(b.cell_start);
        dfree(b.cell_fill);
        const int alloc = std::max(off, b.cell_alloc + b.cell_alloc / 2);
        b.cell_count = dalloc<int>(alloc + 1);
        b.cell_start = dalloc<int>(alloc + 1);
        b.cell_fill = dalloc<int>(std::max(alloc + 1, b.n));
        b.cell_alloc = alloc;
        b.ncells = -1;
        b.alloc_gen++;
    }
    if (off != b.ncells) {
        // each rebuild's scan clears the counts it consumed: zero them when
        // the grid changes
        HIPCHECK(hipMemsetAsync(b.cell_count, 0, sizeof(int) * (b.cell_alloc + 1), b.ctx->stream));
        b.ncells = off;
    }
    HIPCHECK(hipMemcpyAsync(b.d_desc, b.desc.data(), sizeof(SysDesc) * b.nsys, hipMemcpyHostToDevice, b.ctx->stream));
}

// grow_only: keep at least the current capacity (refills of a running
// batch).
void size_list(Batch& b, bool grow_only = false) {
    // initial capacity from the number density of the atoms each box
    // holds: 1.3 x (4/3) pi rns^3 n + 16
    double est = 16;
    for (int s = 0; s < b.nsys; ++s) {
        const SysDesc& d = b.desc[s];
        double vol = 1.0;
        for (int a = 0; a < 3; ++a) vol *= d.L[a] > 0 ? d.L[a] : 1.0;
        double nat = b.present[s];
        double r = b.rns_max[s];
        est = std::max(est, 1.3 * nat / vol * 4.18879 * r * r * r + 16);
    }
    est = std::min(est, 4096.0);
    int cap = (((int)est + b.tpa - 1) / b.tpa) * b.tpa;
    if (grow_only && b.nl && b.cap % b.tpa == 0) cap = std::max(cap, b.cap);
    const int A = kWave / b.tpa;
    size_t waves = (size_t)(b.n + A - 1) / A;
    size_t elems = waves * (size_t)(cap / b.tpa) * kWave;
    if (elems > b.nl_elems || !b.nl) {
        dfree(b.nl);
        b.nl = dalloc<int>(elems);
        b.nl_elems = elems;
        b.alloc_gen++;
    }
    b.cap = cap;
    b.M = cap / b.tpa;
}

void sync_where(Batch& b, std::vector<int>& where) {
    where.resize(b.n);
    HIPCHECK(hipMemcpyAsync(where.data(), b.where, sizeof(int) * b.n, hipMemcpyDeviceToHost, b.ctx->stream));
    HIPCHECK(hipStreamSynchronize(b.ctx->stream));
}

void upload(Batch& b, const AtomRec* a, uint32_t fields) {
    const int n = b.n;
    if (n == 0) return;
    hipStream_t st = b.ctx->stream;
    std::vector<double> stage((size_t)D_NFIELDS * n, 0.0);
    std::vector<double4> spos(n);
    uint32_t dmask = 0;
    auto put = [&](int f, int o, double v) { stage[(size_t)f * n + o] = v; };
    bool rns_changed = false;
    for (int s = 0; s < b.nsys; ++s) {
        for (int o = b.atom_off[s]; o < b.atom_off[s + 1]; ++o) {
            const AtomRec& r = a[o];
            if (fields & MDC_F_KIND) {
                if (r.kind > 255) throw Error(E_INVAL, "atom kind above 255 is not supported");
                b.kind[o] = r.kind;
                b.kind_max = -1;
            }
            if (fields & MDC_F_TAG) b.tag[o] = r.tag;
            if (fields & MDC_F_UID) b.uid[o] = r.uid;
            if (fields & MDC_F_RCUT) b.rcut_host[o] = r.rcut;
            spos[o] = make_double4(r.r[0], r.r[1], r.r[2], pack_w(b.kind[o], b.tag[o] == 0 ? 1 : 0, s));
            if (fields & MDC_F_V)
                for (int c = 0; c < 3; ++c) put(D_VX + c, o, r.v[c]);
            if (fields & MDC_F_F)
                for (int c = 0; c < 3; ++c) put(D_FX + c, o, r.f[c]);
            if (fields & MDC_F_VIR) {
                // column-major V[col*3 + row]
                put(D_VXX, o, r.V[0]);
                put(D_VYY, o, r.V[4]);
                put(D_VZZ, o, r.V[8]);
                put(D_VXY, o, r.V[3]);
                put(D_VXZ, o, r.V[6]);
                put(D_VYZ, o, r.V[7]);
            }
            if (fields & MDC_F_M) put(D_M, o, r.m);
            if (fields & MDC_F_RNS) {
                if (!(r.rns >= 0) || !std::isfinite(r.rns)) throw Error(E_INVAL, "rns must be finite and >= 0");
                put(D_RNS, o, r.rns);
            }
            if (fields & MDC_F_EP) put(D_EP, o, r.Ep);
            if (fields & MDC_F_EM) put(D_EM, o, r.Em);
            if (fields & MDC_F_NC) put(D_NC, o, r.nc);
        }
    }
    if (fields & MDC_F_V) dmask |= 7u << D_VX;
    if (fields & MDC_F_F) dmask |= 7u << D_FX;
    if (fields & MDC_F_VIR) dmask |= 63u << D_VXX;
    if (fields & MDC_F_M) dmask |= 1u << D_M;
    if (fields & MDC_F_RNS) dmask |= 1u << D_RNS;
    if (fields & MDC_F_EP) dmask |= 1u << D_EP;
    if (fields & MDC_F_EM) dmask |= 1u << D_EM;
    if (fields & MDC_F_NC) dmask |= 1u << D_NC;
    if (fields & MDC_F_RNS) {
        for (int s = 0; s < b.nsys; ++s) {
            double mn = 1e300, mx = 0;
            for (int o = b.atom_off[s]; o < b.atom_off[s + 1]; ++o) {
                mn = std::min(mn, a[o].rns);
                mx = std::max(mx, a[o].rns);
            }
            if (b.atom_off[s + 1] == b.atom_off[s]) mn = mx = 0;
            if (mn != b.rns_min[s] || mx != b.rns_max[s]) rns_changed = true;
            b.rns_min[s] = mn;
            b.rns_max[s] = mx;
        }
    }
    const size_t sbytes = stage.size() * sizeof(double);
    char* area = stage_area(b, sbytes + spos.size() * sizeof(double4));
    double* d_stage = reinterpret_cast<double*>(area);
    double4* d_spos = reinterpret_cast<double4*>(area + sbytes);
    HIPCHECK(hipMemcpyAsync(d_stage, stage.data(), sbytes, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(d_spos, spos.data(), spos.size() * sizeof(double4), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_scatter_in, dim3((n + 255) / 256), dim3(256), 0, st, d_stage, d_spos, b.where, b.pos[b.cur],
                       b.dst[b.sc], b.fst[b.sc], dmask, (fields & MDC_F_R) ? 1 : 0, (fields & MDC_F_F) ? 1 : 0, n);
    HIPCHECK(hipStreamSynchronize(st));
    if (fields & MDC_F_F) b.f_prec = P_FP64;
    if (fields & MDC_F_RCUT) b.last_rcut = -1;
    if (fields & MDC_F_R) {
        b.shifts_stale = true;
        reset_rmin(b);
    }
    // every record now holds an atom (the tag words were written afresh):
    // records of the void tail became live, so the next build sorts all
    long long held = 0;
    for (int s = 0; s < b.nsys; ++s) held += b.present[s];
    if (held < n) {
        b.full_sort = true;
        b.list_valid = false;
    }
    for (int s = 0; s < b.nsys; ++s) b.present[s] = b.atom_off[s + 1] - b.atom_off[s];
    if (rns_changed) {
        layout_cells(b);
        size_list(b);
        b.list_valid = false;
    }
}

// Box `box` takes `count` records of the caller (at most its record count)
// in a new geometry; its remaining records become void (zero fields, tag
// 0, binned into the box's void cell).  Positions, velocities, forces and
// every other field come from the records.  The list is out of date until
// the next build; the other boxes keep their state.
void refill_box(Batch& b, int box, int count, const AtomRec* a, const double* lo, const double* hi,
                const uint8_t* per) {
    if (box < 0 || box >= b.nsys) throw Error(E_INVAL, "box index out of range");
    const int first = b.atom_off[box], cap = b.atom_off[box + 1] - first;
    if (count < 0 || count > cap) throw Error(E_INVAL, "more atoms than the box's records");
    if (count > 0 && !a) throw Error(E_INVAL, "null atom records");
    SysDesc d = b.desc[box];
    d.any_per = 0;
    for (int ax = 0; ax < 3; ++ax) {
        d.lo[ax] = lo[ax];
        d.hi[ax] = hi[ax];
        d.L[ax] = hi[ax] - lo[ax];
        d.per[ax] = per[ax] ? 1 : 0;
        if (d.per[ax] && !(d.L[ax] > 0)) throw Error(E_INVAL, "periodic axis with non-positive length");
        d.any_per |= (d.per[ax] && b.periodic_bc[box]) ? 1 : 0;
    }
    double mn = 1e300, mx = 0.0;
    for (int q = 0; q < count; ++q) {
        const AtomRec& r = a[q];
        if (r.kind > 255) throw Error(E_INVAL, "atom kind above 255 is not supported");
        if (!(r.rns >= 0) || !std::isfinite(r.rns)) throw Error(E_INVAL, "rns must be finite and >= 0");
        mn = std::min(mn, r.rns);
        mx = std::max(mx, r.rns);
    }
    if (count == 0) mn = mx = 0.0;
    if (cap == 0) return;
    // the atoms staged with stride count; the void records are written on
    // the device
    std::vector<double> stage((size_t)D_NFIELDS * count, 0.0);
    std::vector<double4> spos(count);
    auto put = [&](int f, int q, double v) { stage[(size_t)f * count + q] = v; };
    for (int q = 0; q < cap; ++q) {
        const int o = first + q;
        if (q < count) {
            const AtomRec& r = a[q];
            b.kind[o] = r.kind;
            b.tag[o] = r.tag;
            b.uid[o] = r.uid;
            b.rcut_host[o] = r.rcut;
            spos[q] = make_double4(r.r[0], r.r[1], r.r[2], pack_w(r.kind, r.tag == 0 ? 1 : 0, box));
            for (int c = 0; c < 3; ++c) {
                put(D_VX + c, q, r.v[c]);
                put(D_FX + c, q, r.f[c]);
            }
            put(D_VXX, q, r.V[0]);
            put(D_VYY, q, r.V[4]);
            put(D_VZZ, q, r.V[8]);
            put(D_VXY, q, r.V[3]);
            put(D_VXZ, q, r.V[6]);
            put(D_VYZ, q, r.V[7]);
            put(D_M, q, r.m);
            put(D_RNS, q, r.rns);
            put(D_EP, q, r.Ep);
            put(D_EM, q, r.Em);
            put(D_NC, q, r.nc);
        } else {
            b.kind[o] = 0;
            b.tag[o] = 0;
            b.uid[o] = 0;
            b.rcut_host[o] = 0.0;
        }
    }
    hipStream_t st = b.ctx->stream;
    const size_t sbytes = stage.size() * sizeof(double);
    char* area = stage_area(b, sbytes + spos.size() * sizeof(double4));
    double* d_stage = reinterpret_cast<double*>(area);
    double4* d_spos = reinterpret_cast<double4*>(area + sbytes);
    if (count > 0) {
        HIPCHECK(hipMemcpyAsync(d_stage, stage.data(), sbytes, hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemcpyAsync(d_spos, spos.data(), spos.size() * sizeof(double4), hipMemcpyHostToDevice, st));
    }
    const double4 void_pos = make_double4(d.lo[0], d.lo[1], d.lo[2], pack_w(0, 1, box, 1));
    hipLaunchKernelGGL(k_scatter_range, dim3((cap + 255) / 256), dim3(256), 0, st, d_stage, d_spos, b.where,
                       b.pos[b.cur], b.dst[b.sc], b.fst[b.sc], first, count, cap, b.n, void_pos);
    HIPCHECK(hipStreamSynchronize(st));
    b.desc[box] = d;
    b.rns_min[box] = mn;
    b.rns_max[box] = mx;
    b.present[box] = count;
    b.kind_max = -1;
    layout_cells(b);
    // the list capacity only grows here: refills come and go in a running
    // batch, and a smaller list would be regrown by the next denser box
    size_list(b, true);
    b.list_valid = false;
    b.full_sort = true;
    b.shifts_stale = true;
    b.last_rcut = -1;
    reset_rmin(b);
}

// records [first, first + count) of the caller's order into a[0, count)
void download(Batch& b, AtomRec* a, uint32_t fields, int first, int count) {
    const int n = b.n;
    if (count <= 0) return;
    hipStream_t st = b.ctx->stream;
    std::vector<double> stage((size_t)D_NFIELDS * count);
    std::vector<double4> spos(count);
    const size_t sbytes = stage.size() * sizeof(double);
    char* area = stage_area(b, sbytes + spos.size() * sizeof(double4));
    double* d_stage = reinterpret_cast<double*>(area);
    double4* d_spos = reinterpret_cast<double4*>(area + sbytes);
    hipLaunchKernelGGL(k_gather_out, dim3((count + 255) / 256), dim3(256), 0, st, d_stage, d_spos, b.where,
                       b.pos[b.cur], b.dst[b.sc], b.fst[b.sc], b.f_prec == P_MIXED ? 1 : 0, n, first, count);
    HIPCHECK(hipMemcpyAsync(stage.data(), d_stage, sbytes, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(spos.data(), d_spos, spos.size() * sizeof(double4), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    auto get = [&](int f, int o) { return stage[(size_t)f * count + (o - first)]; };
    for (int s = 0; s < b.nsys; ++s) {
        const SysDesc& d = b.desc[s];
        const int o0 = std::max(first, b.atom_off[s]), o1 = std::min(first + count, b.atom_off[s + 1]);
        for (int o = o0; o < o1; ++o) {
            AtomRec& r = a[o - first];
            if (fields & MDC_F_R) {
                r.r[0] = spos[o - first].x;
                r.r[1] = spos[o - first].y;
                r.r[2] = spos[o - first].z;
            }
            if (fields & MDC_F_V)
                for (int c = 0; c < 3; ++c) r.v[c] = get(D_VX + c, o);
            if (fields & MDC_F_F)
                for (int c = 0; c < 3; ++c) r.f[c] = get(D_FX + c, o);
            if (fields & MDC_F_VIR) {
                double xx = get(D_VXX, o), yy = get(D_VYY, o), zz = get(D_VZZ, o);
                double xy = get(D_VXY, o), xz = get(D_VXZ, o), yz = get(D_VYZ, o);
                double V[9] = {xx, xy, xz, xy, yy, yz, xz, yz, zz};
                std::memcpy(r.V, V, sizeof(V));
            }
            if (fields & MDC_F_M) r.m = get(D_M, o);
            if (fields & MDC_F_RCUT) r.rcut = b.last_rcut >= 0 && b.tag[o] != 0 ? b.last_rcut : b.rcut_host[o];
            if (fields & MDC_F_RNS) r.rns = get(D_RNS, o);
            if (fields & MDC_F_EP) r.Ep = get(D_EP, o);
            if (fields & MDC_F_EM) r.Em = get(D_EM, o);
            if (fields & MDC_F_NC) r.nc = get(D_NC, o);
            if (fields & MDC_F_KIND) r.kind = b.kind[o];
            if (fields & MDC_F_UID) r.uid = b.uid[o];
            if (fields & MDC_F_TAG) r.tag = b.tag[o];
            if (fields & MDC_F_BC) {
                // boundary::Periodic::need_apply (periodic.cxx:9-21)
                bool need = false;
                if (b.periodic_bc[s] && b.tag[o] != 0) {
                    const double rr[3] = {spos[o - first].x, spos[o - first].y, spos[o - first].z};
                    for (int ax = 0; ax < 3; ++ax)
                        if (d.per[ax]) {
                            double dist = std::min(std::fabs(rr[ax] - d.lo[ax]), std::fabs(rr[ax] - d.hi[ax]));
                            need = need || dist < get(D_RNS, o);
                        }
                }
                r.bc = need ? 1 : 0;
            }
        }
    }
}

// VerletList::optimize_search_radius (verlet-list.cxx:71-92): rns_i =
// kbuf * the largest rcut_i + rcut_j over atom i's current list, self
// included.  rc holds rcut in list (slot) order.
__global__ void k_opt_rns(const int* __restrict__ nl, const int* __restrict__ nn, const double* __restrict__ rc,
                          double* __restrict__ rns, int tpa, int M, int cap, int n, double kbuf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double ri = rc[i];
    double h = ri + ri;
    const int c = min(nn[i], cap);
    for (int k = 0; k < c; ++k) h = fmax(h, ri + rc[ent_j(nl[list_index(i, k, tpa, M)])]);
    rns[i] = kbuf * h;
}

// New rns of every atom from the current list (zero for an atom with no
// list yet, as the reference's empty lists give), in the caller's order;
// the list is then out of date and the cell layout follows the new radii.
void optimize_rns(Batch& b, double kbuf, double* rns_out) {
    const int n = b.n;
    if (n == 0) return;
    hipStream_t st = b.ctx->stream;
    std::vector<double> rns(n, 0.0);
    if (b.list_valid) {
        std::vector<int> where;
        sync_where(b, where);
        std::vector<double> rc(n);
        for (int o = 0; o < n; ++o) rc[where[o]] = b.rcut_host[o];
        double* d_rc = dalloc<double>(n);
        try {
            HIPCHECK(hipMemcpyAsync(d_rc, rc.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
            // the live records (the void tail keeps its zero radii)
            if (b.live > 0)
                hipLaunchKernelGGL(k_opt_rns, dim3((b.live + 255) / 256), dim3(256), 0, st, b.nl, b.nn, d_rc,
                                   b.d(b.sc, D_RNS), b.tpa, b.M, b.cap, b.live, kbuf);
            HIPCHECK(hipMemcpyAsync(rc.data(), b.d(b.sc, D_RNS), sizeof(double) * n, hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
        } catch (...) {
            dfree(d_rc);
            throw;
        }
        dfree(d_rc);
        for (int o = 0; o < n; ++o) rns[o] = rc[where[o]];
    } else {
        HIPCHECK(hipMemsetAsync(b.d(b.sc, D_RNS), 0, sizeof(double) * n, st));
        HIPCHECK(hipStreamSynchronize(st));
    }
    for (int s = 0; s < b.nsys; ++s) {
        // over the atoms the box holds
        double mn = 1e300, mx = 0;
        for (int o = b.atom_off[s]; o < b.atom_off[s] + b.present[s]; ++o) {
            mn = std::min(mn, rns[o]);
            mx = std::max(mx, rns[o]);
        }
        if (b.present[s] == 0) mn = mx = 0;
        b.rns_min[s] = mn;
        b.rns_max[s] = mx;
    }
    layout_cells(b);
    size_list(b);
    b.list_valid = false;
    if (rns_out) std::memcpy(rns_out, rns.data(), sizeof(double) * n);
}

void check_kinds(Batch& b, const Pot& p) {
    const int limit = p.type == PT_BPNN ? bpnn_species(p) : (p.type == PT_EAM && p.eam.nk > 1 ? p.eam.nk : 0);
    if (limit == 0) return;
    if (b.kind_max < 0) {
        // recounted once after kinds change, not per evaluation
        int mx = 0;
        for (int o = 0; o < b.n; ++o) mx = std::max(mx, (int)b.kind[o]);
        b.kind_max = mx;
    }
    if (b.kind_max >= limit)
        throw Error(E_INVAL, p.type == PT_BPNN ? "atom kind has no network in the MAISE model"
                             : p.eam.alloy  ? "atom kind outside the alloy's elements"
                                            : "atom kind outside the Multi potential's kinds");
}

}  // namespace

namespace mdc {
int report_error(Ctx* c, int code, const std::string& msg) {
    g_last = msg;
    if (c) c->err = msg;
    return code;
}
}  // namespace mdc

extern "C" {

const char* mdc_last_error(void) { return g_last.c_str(); }
const char* mdc_ctx_last_error(const mdc_ctx* ctx) { return ctx ? ctx->c.err.c_str() : g_last.c_str(); }
int mdc_atom_record_size(void) { return (int)sizeof(AtomRec); }

mdc_status mdc_ctx_create(int device, mdc_ctx** out) {
    return guard(nullptr, [&] {
        if (!out) throw Error(E_INVAL, "null output pointer");
        int ndev = 0;
        HIPCHECK(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) throw Error(E_INVAL, "no such device");
        HIPCHECK(hipSetDevice(device));
        auto c = std::make_unique<mdc_ctx>();
        c->c.device = device;
        HIPCHECK(hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking));
        void* pin = nullptr;
        HIPCHECK(hipHostMalloc(&pin, sizeof(int) * 3 * M_COUNT, hipHostMallocDefault));
        c->c.h_pin = static_cast<int*>(pin);
        std::memset(c->c.h_pin, 0, sizeof(int) * 3 * M_COUNT);
        for (auto& ev : c->c.chunk_ev) HIPCHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        hipDeviceProp_t prop;
        HIPCHECK(hipGetDeviceProperties(&prop, device));
        c->c.cu_count = prop.multiProcessorCount;
        c->c.lds_max = prop.sharedMemPerBlock;
        scratch_lookup(device);
        *out = c.release();
    });
}

mdc_status mdc_ctx_set_spin(mdc_ctx* ctx, int spin) {
    return guard(ctx ? &ctx->c : nullptr, [&] {
        if (!ctx) throw Error(E_INVAL, "null argument");
        ctx->c.spin = spin != 0;
    });
}

mdc_status mdc_ctx_destroy(mdc_ctx* ctx) {
    if (!ctx) return 0;
    (void)hipSetDevice(ctx->c.device);
    if (ctx->c.stream) (void)hipStreamDestroy(ctx->c.stream);
    if (ctx->c.h_pin) (void)hipHostFree(ctx->c.h_pin);
    for (auto ev : ctx->c.chunk_ev)
        if (ev) (void)hipEventDestroy(ev);
    delete ctx;
    return 0;
}

mdc_status mdc_pot_lj_create(mdc_ctx* ctx, double a, double s, double rcut, mdc_pot** out) {
    return guard(ctx ? &ctx->c : nullptr, [&] {
        if (!ctx || !out) throw Error(E_INVAL, "null argument");
        if (!(rcut > 0) || !(s > 0)) throw Error(E_INVAL, "LJ sigma and cutoff must be positive");
        auto p = std::make_unique<Pot>();
        p->ctx = &ctx->c;
        p->type = PT_LJ;
        p->rcut = rcut;
        // LJs-cut.cxx:8-20
        LJParams& L = p->lj;
        L.a = a;
        L.s = s;
        L.s6 = std::pow(s, 6);
        L.lj1 = 2 * L.s6;
        L.lj2 = 1.0;
        L.factor = 48.0 * a * L.s6;
        L.rc = rcut;
        L.rc2 = rcut * rcut;
        *out = reinterpret_cast<mdc_pot*>(p.release());
    });
}

mdc_status mdc_pot_ljs_create(mdc_ctx* ctx, double aVr, double rVr, double rcut, mdc_pot** out) {
    return guard(ctx ? &ctx->c : nullptr, [&] {
        if (!ctx || !out) throw Error(E_INVAL, "null argument");
        if (!(rcut > 0) || !(rVr > 0)) throw Error(E_INVAL, "LJs rVr and cutoff must be positive");
        auto p = std::make_unique<Pot>();
        p->ctx = &ctx->c;
        p->type = PT_LJ;
        p->rcut = rcut;
        LJParams& L = p->lj;
        L.fn = 1;
        L.rc = rcut;
        L.rc2 = rcut * rcut;
        // LJs.cxx:7-26, in the constructor's order of operations
        const double X2min = 1.2599210498948732;  // 2^(1/3): X = 0 at the LJ minimum
        const double R2cut = L.rc2;
        double a = aVr;
        const double r2v = rVr * rVr;
        L.rVr2 = r2v;
        L.rVr1 = 1.0 / r2v;
        double aVr2 = R2cut / r2v - X2min;
        double aLJ3 = (r2v / R2cut) * (r2v / R2cut) * (r2v / R2cut);
        double bLJ2 = (aLJ3 - 0.5) * aVr2 * r2v / R2cut;
        L.aLJ = (aLJ3 - 1.0 + 2.0 * bLJ2) * aLJ3 / (aVr2 * aVr2);
        L.bLJ = -(aLJ3 - 1.0 + 3.0 * bLJ2) * aLJ3 / (aVr2 * aVr2 * aVr2);
        L.aLJ3 = 3.0 * L.aLJ;
        L.bLJ2 = 2.0 * L.bLJ;
        L.aVr2 = -48.0 * a / r2v;
        L.aVr = 2.0 * a;  // V(r)/2 per pair
        *out = reinterpret_cast<mdc_pot*>(p.release());
    });
}

// Tabulated Pair (pair.cxx:15-34): r*phi on nr knots of step dr, spline
// slopes as EAM's with the ib = 1 end condition.
static Pot* make_pair(Ctx* ctx, double rcut, const double* rphi, int nr, double dr) {
    if (nr < 3) throw Error(E_INVAL, "Pair tables need at least 3 knots");
    if (!(dr > 0) || !(rcut > 0)) throw Error(E_INVAL, "Pair step and cutoff must be positive");
    auto p = std::make_unique<Pot>();
    p->ctx = ctx;
    p->type = PT_LJ;
    p->rcut = rcut;
    p->lj.fn = 2;
    p->lj.rc = rcut;
    p->lj.rc2 = rcut * rcut;
    EamTables& T = p->eam;
    T.nr = nr;
    T.dr = dr;
    T.rc = rcut;
    std::vector<double> rec64, z(nr);
    spline_slopes(rphi, z.data(), 1.0 / dr, nr, 1);
    records(rphi, z.data(), nr, rec64);
    for (double x : rec64)
        if (!std::isfinite(x)) throw Error(E_RANGE, "non-finite value in the Pair table");
    std::vector<float> rec32(rec64.size());
    for (size_t q = 0; q < rec64.size(); q += 4) {
        const double y0 = rec64[q], z0 = rec64[q + 1], y1 = rec64[q + 2], z1 = rec64[q + 3];
        const double sl = (y1 - y0) / dr;
        const double aa = (z1 - sl) - (sl - z0);
        const double bb = aa - (sl - z0);
        rec32[q] = (float)y0;
        rec32[q + 1] = (float)(dr * z0);
        rec32[q + 2] = (float)(-dr * bb);
        rec32[q + 3] = (float)(dr * aa);
    }
    T.nrec = rec64.size() / 4;
    T.r64 = dalloc<double4>(T.nrec);
    T.r32 = dalloc<float4>(T.nrec);
    HIPCHECK(hipMemcpy(T.r64, rec64.data(), rec64.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(T.r32, rec32.data(), rec32.size() * sizeof(float), hipMemcpyHostToDevice));
    return p.release();
}

mdc_status mdc_pot_pair_create(mdc_ctx* ctx, double rcut, const double* rphi, int nr, double dr, mdc_pot** out) {
    return guard(ctx ? &ctx->c : nullptr, [&] {
        if (!ctx || !out || !rphi) throw Error(E_INVAL, "null argument");
        *out = reinterpret_cast<mdc_pot*>(make_pair(&ctx->c, rcut, rphi, nr, dr));
    });
}

// Pair(filename) (pair.cxx:36-56): four lines skipped, then "Nr dr rcut"
// in Angstrom and Nr values of r*phi in eV*Angstrom
mdc_status mdc_pot_pair_create_file(mdc_ctx* ctx, const char* path, mdc_pot** out) {
    return guard(ctx ? &ctx->c : nullptr, [&] {
        if (!ctx || !out || !path) throw Error(E_INVAL, "null argument");
        std::ifstream f(path);
        if (!f.is_open()) throw Error(E_FILE, std::string("Failed to open file: ") + path);
        std::string line;
        for (int i = 0; i < 4; ++i) std::getline(f, line);
        int nr = 0;
        double dr = 0, rc = 0;
        f >> nr >> dr >> rc;
        if (!f || nr < 3) throw Error(E_FILE, std::string("malformed Pair header: ") + path);
        const double Na = 6.02214076e+23, Qe = 1.602176634e-19;
        const double UnMass = 1e-3 / Na, UnLeng = 1e-9, UnTime = 1e-12;
        const double UnEnrg = UnMass * (UnLeng / UnTime) * (UnLeng / UnTime);
        std::vector<double> rphi(nr);
        for (auto& x : rphi) {
            f >> x;
            x *= Qe / (10 * UnEnrg);
        }
        if (!f) throw Error(E_FILE, std::string("truncated Pair table: ") + path);
        *out = reinterpret_cast<mdc_pot*>(make_pair(&ctx->c, rc / 10, rphi.data(), nr, dr / 10));
    });
}

mdc_status mdc_pot_eam_create(mdc_ctx* ctx, double rcut, const double* F, const double* rho, const double* rphi,
                              int nr, double dr, int nrho, double drho, mdc_pot** out) {
    return guard(ctx ? &ctx->c : nullptr, [&] {
        if (!ctx || !out || !F || !rho || !rphi) throw Error(E_INVAL, "null argument");
        HIPCHECK(hipSetDevice(ctx->c.device));
        *out = reinterpret_cast<mdc_pot*>(make_eam(&ctx->c, 1, rcut, F, rho, rphi, nr, dr, nrho, drho));
    });
}

mdc_status mdc_pot_multi_create(mdc_ctx* ctx, int nkinds, int npots, const int* kind_i, const int* kind_j,
                                const mdc_pot* const* pots, mdc_pot** out) {
    return guard(ctx ? &ctx->c : nullptr, [&] {
        if (!ctx || !out || !kind_i || !kind_j || !pots || npots < 1) throw Error(E_INVAL, "null argument");
        HIPCHECK(hipSetDevice(ctx->c.device));
        std::vector<const Pot*> pp(npots);
        for (int q = 0; q < npots; ++q) pp[q] = pots[q] ? &reinterpret_cast<const mdc_pot*>(pots[q])->p : nullptr;
        *out = reinterpret_cast<mdc_pot*>(make_multi(&ctx->c, nkinds, npots, kind_i, kind_j, pp.data()));
    });
}

mdc_status mdc_pot_eam_create_file(mdc_ctx* ctx, const char* path, mdc_pot** out) {
    return guard(ctx ? &ctx->c : nullptr, [&] {
        if (!ctx || !out || !path) throw Error(E_INVAL, "null argument");
        HIPCHECK(hipSetDevice(ctx->c.device));
        *out = reinterpret_cast<mdc_pot*>(make_eam_file(&ctx->c, path));
    });
}

mdc_status mdc_pot_alloy_create(mdc_ctx* ctx, int nelements, double rcut, const double* F, const double* rho,
                                const double* rphi, int nr, double dr, int nrho, double drho, mdc_pot** out) {
    return guard(ctx ? &ctx->c : nullptr, [&] {
        if (!ctx || !out || !F || !rho || !rphi) throw Error(E_INVAL, "null argument");
        HIPCHECK(hipSetDevice(ctx->c.device));
        *out = reinterpret_cast<mdc_pot*>(make_eam(&ctx->c, nelements, rcut, F, rho, rphi, nr, dr, nrho, drho));
    });
}

mdc_status mdc_pot_alloy_create_file(mdc_ctx* ctx, const char* path, mdc_pot** out) {
    return guard(ctx ? &ctx->c : nullptr, [&] {
        if (!ctx || !out || !path) throw Error(E_INVAL, "null argument");
        HIPCHECK(hipSetDevice(ctx->c.device));
        *out = reinterpret_cast<mdc_pot*>(make_alloy_file(&ctx->c, path));
    });
}

mdc_status mdc_pot_bpnn_create(mdc_ctx* ctx, const char* path, mdc_pot** out) {
    return guard(ctx ? &ctx->c : nullptr, [&] {
        if (!ctx || !out || !path) throw Error(E_INVAL, "null argument");
        HIPCHECK(hipSetDevice(ctx->c.device));
        *out = reinterpret_cast<mdc_pot*>(bpnn_create(&ctx->c, path));
    });
}

mdc_status mdc_pot_bpnn_describe(const char* maise_path, int* arch, int* counts, double* rcut) {
    return guard(nullptr, [&] {
        if (!maise_path) throw Error(E_INVAL, "null path");
        bpnn_describe(maise_path, arch, counts, rcut);
    });
}

mdc_status mdc_pot_rcut(const mdc_pot* pot, double* rcut) {
    if (!pot || !rcut) return E_INVAL;
    *rcut = pot->p.rcut;
    return 0;
}

mdc_status mdc_pot_destroy(mdc_pot* pot) {
    if (!pot) return 0;
    (void)hipSetDevice(pot->p.ctx->device);
    free_pot(&pot->p);
    return 0;
}

mdc_status mdc_sys_create(mdc_ctx* ctx, int nsys, const uint64_t* natoms, const double* lo, const double* hi,
                          const uint8_t* periodic, int periodic_boundary, mdc_sys** out) {
    return guard(ctx ? &ctx->c : nullptr, [&] {
        if (!ctx || !out || nsys < 1 || !natoms || !lo || !hi || !periodic) throw Error(E_INVAL, "bad argument");
        if (nsys > (1 << 20)) throw Error(E_INVAL, "too many systems in one batch");
        HIPCHECK(hipSetDevice(ctx->c.device));
        auto s = std::make_unique<mdc_sys>();
        Batch& b = s->b;
        b.ctx = &ctx->c;
        b.nsys = nsys;
        b.atom_off.assign(nsys + 1, 0);
        uint64_t total = 0;
        for (int i = 0; i < nsys; ++i) {
            // checked as 64-bit counts before any narrowing
            if (natoms[i] >= (uint64_t)kMaxAtomsShift) throw Error(E_INVAL, "at most 2^26 - 1 atoms per batch");
            total += natoms[i];
            if (total >= (uint64_t)kMaxAtomsShift) throw Error(E_INVAL, "at most 2^26 - 1 atoms per batch");
            b.atom_off[i + 1] = (int)total;
        }
        b.n = b.atom_off[nsys];
        if (b.n >= kMaxAtomsShift) throw Error(E_INVAL, "at most 2^26 - 1 atoms per batch");
        b.desc.resize(nsys);
        b.periodic_bc.assign(nsys, periodic_boundary ? 1 : 0);
        for (int i = 0; i < nsys; ++i) {
            SysDesc& d = b.desc[i];
            std::memset(&d, 0, sizeof(d));
            d.any_per = 0;
            for (int a = 0; a < 3; ++a) {
                d.lo[a] = lo[3 * i + a];
                d.hi[a] = hi[3 * i + a];
                d.L[a] = d.hi[a] - d.lo[a];
                // the image rule of the ghost pass applies only when the solver
                // carries the Periodic boundary
                d.per[a] = periodic[3 * i + a] ? 1 : 0;
                if (d.per[a] && !(d.L[a] > 0)) throw Error(E_INVAL, "periodic axis with non-positive length");
                d.any_per |= (d.per[a] && periodic_boundary) ? 1 : 0;
            }
            d.atom_off = b.atom_off[i];
            d.natoms = b.atom_off[i + 1] - b.atom_off[i];
            d.trig2 = -1.0;
        }
        b.rns_min.assign(nsys, 0.0);
        b.rns_max.assign(nsys, 0.0);
        b.present.resize(nsys);
        for (int i = 0; i < nsys; ++i) b.present[i] = b.atom_off[i + 1] - b.atom_off[i];
        b.live = b.n;
        b.full_sort = true;
        b.kind.assign(b.n, 0);
        b.tag.assign(b.n, 65535);
        b.uid.assign(b.n, 0);
        b.rcut_host.assign(b.n, 0.0);
        const int n = b.n;
        try {
            b.d_desc = dalloc<SysDesc>(nsys);
            b.pos[0] = dalloc<double4>(n);
            b.pos[1] = dalloc<double4>(n);
            for (int k = 0; k < 2; ++k) {
                // at least one record each, as the clears below assume
                b.dst[k] = dalloc<double>((size_t)D_NFIELDS * std::max(n, 1));
                b.fst[k] = dalloc<float>((size_t)F_NFIELDS * std::max(n, 1));
                b.orig[k] = dalloc<int>(n);
                HIPCHECK(hipMemset(b.dst[k], 0, sizeof(double) * D_NFIELDS * std::max(n, 1)));
                HIPCHECK(hipMemset(b.fst[k], 0, sizeof(float) * F_NFIELDS * std::max(n, 1)));
            }
            b.rsave = dalloc<double4>(n);
            b.posF = dalloc<double4>(n);
            b.where = dalloc<int>(n);
            b.atom_cell = dalloc<int>(n);
            b.sorted_src = dalloc<int>(n);
            b.nn = dalloc<int>(n);
            b.misc = dalloc<int>(M_COUNT);
            b.thermo = dalloc<double>((size_t)9 * nsys + 2);
            HIPCHECK(hipMemset(b.misc, 0, sizeof(int) * M_COUNT));
            const int rmin_none = kRminNone;
            HIPCHECK(hipMemcpy(b.misc + M_RMIN, &rmin_none, sizeof(int), hipMemcpyHostToDevice));
            HIPCHECK(hipMemset(b.pos[0], 0, sizeof(double4) * std::max(n, 1)));
            HIPCHECK(hipMemset(b.rsave, 0, sizeof(double4) * std::max(n, 1)));
            HIPCHECK(hipMemset(b.nn, 0, sizeof(int) * std::max(n, 1)));
            std::vector<int> iota(n);
            std::iota(iota.begin(), iota.end(), 0);
            HIPCHECK(hipMemcpy(b.where, iota.data(), sizeof(int) * n, hipMemcpyHostToDevice));
            HIPCHECK(hipMemcpy(b.orig[0], iota.data(), sizeof(int) * n, hipMemcpyHostToDevice));
            b.tpa = n >= 200000 ? 4 : 8;
            // the staged tables pay off once a launch has thousands of
            // wavefronts (mdc_sys_autotune measures it)
            b.q_lds = n >= 65536;
            layout_cells(b);
            size_list(b);
        } catch (...) {
            free_batch(b);
            throw;
        }
        *out = s.release();
    });
}

mdc_status mdc_sys_destroy(mdc_sys* sys) {
    if (!sys) return 0;
    (void)hipSetDevice(sys->b.ctx->device);
    (void)hipStreamSynchronize(sys->b.ctx->stream);
    free_batch(sys->b);
    delete sys;
    return 0;
}

mdc_status mdc_sys_upload(mdc_sys* sys, const void* atoms, uint32_t fields) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys || !atoms) throw Error(E_INVAL, "null argument");
        HIPCHECK(hipSetDevice(sys->b.ctx->device));
        upload(sys->b, static_cast<const AtomRec*>(atoms), fields);
    });
}

mdc_status mdc_sys_download(mdc_sys* sys, void* atoms, uint32_t fields) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys || !atoms) throw Error(E_INVAL, "null argument");
        HIPCHECK(hipSetDevice(sys->b.ctx->device));
        download(sys->b, static_cast<AtomRec*>(atoms), fields, 0, sys->b.n);
    });
}

mdc_status mdc_sys_download_range(mdc_sys* sys, uint64_t first, uint64_t count, void* atoms, uint32_t fields) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys || (!atoms && count)) throw Error(E_INVAL, "null argument");
        if (first > (uint64_t)sys->b.n || count > (uint64_t)sys->b.n - first)
            throw Error(E_INVAL, "record range outside the system");
        HIPCHECK(hipSetDevice(sys->b.ctx->device));
        download(sys->b, static_cast<AtomRec*>(atoms), fields, (int)first, (int)count);
    });
}

mdc_status mdc_sys_set_active(mdc_sys* sys, int box, int active) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys) throw Error(E_INVAL, "null argument");
        Batch& b = sys->b;
        if (box < 0 || box >= b.nsys) throw Error(E_INVAL, "box index out of range");
        HIPCHECK(hipSetDevice(b.ctx->device));
        const int first = b.atom_off[box], count = b.present[box];
        if (count == 0) return;
        // the skip flag travels with the positions; atoms whose own tag is 0
        // stay skipped
        bool any_tag0 = false;
        for (int o = first; o < first + count; ++o) any_tag0 = any_tag0 || b.tag[o] == 0;
        if (active && any_tag0) throw Error(E_INVAL, "box holds atoms with tag 0; re-upload it to reactivate");
        hipLaunchKernelGGL(k_set_inactive, dim3((count + 255) / 256), dim3(256), 0, b.ctx->stream, b.pos[b.cur],
                           b.where, first, count, active ? 0 : 1);
        HIPCHECK(hipGetLastError());
    });
}

mdc_status mdc_sys_refill_box(mdc_sys* sys, int box, uint64_t count, const void* atoms, const double* box_lo,
                              const double* box_hi, const uint8_t* periodic) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys || !box_lo || !box_hi || !periodic) throw Error(E_INVAL, "null argument");
        if (count > (uint64_t)kMaxAtomsShift) throw Error(E_INVAL, "more atoms than the box's records");
        HIPCHECK(hipSetDevice(sys->b.ctx->device));
        refill_box(sys->b, box, (int)count, static_cast<const AtomRec*>(atoms), box_lo, box_hi, periodic);
    });
}

mdc_status mdc_sys_natoms(const mdc_sys* sys, uint64_t* n) {
    if (!sys || !n) return E_INVAL;
    *n = (uint64_t)sys->b.n;
    return 0;
}

static void set_lanes(Batch& b, int lanes) {
    if (lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8 && lanes != 16)
        throw Error(E_INVAL, "lanes per atom must be 1, 2, 4, 8 or 16");
    if (lanes == b.tpa) return;
    HIPCHECK(hipSetDevice(b.ctx->device));
    int maxcap = b.cap;
    b.tpa = lanes;
    size_list(b);
    if (maxcap > b.cap) {
        // keep the grown capacity
        int cap = ((maxcap + b.tpa - 1) / b.tpa) * b.tpa;
        const int A = kWave / b.tpa;
        size_t elems = (size_t)((b.n + A - 1) / A) * (size_t)(cap / b.tpa) * kWave;
        if (elems > b.nl_elems) {
            dfree(b.nl);
            b.nl = dalloc<int>(elems);
            b.nl_elems = elems;
            b.alloc_gen++;
        }
        b.cap = cap;
        b.M = cap / b.tpa;
    }
    b.list_valid = false;
}

mdc_status mdc_sys_set_lanes(mdc_sys* sys, int lanes) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys) throw Error(E_INVAL, "null argument");
        set_lanes(sys->b, lanes);
    });
}

mdc_status mdc_nlist_build(mdc_sys* sys) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys) throw Error(E_INVAL, "null argument");
        HIPCHECK(hipSetDevice(sys->b.ctx->device));
        hipEvent_t t0 = timer_begin(sys->b);
        nlist_rebuild(sys->b);
        timer_end(sys->b, t0, K_REBUILD);
    });
}

mdc_status mdc_nlist_decide(mdc_sys* sys, double rcut, int* need) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys || !need) throw Error(E_INVAL, "null argument");
        HIPCHECK(hipSetDevice(sys->b.ctx->device));
        *need = nlist_decide(sys->b, rcut);
    });
}

mdc_status mdc_nlist_optimize_rns(mdc_sys* sys, double kbuf, double* rns_out) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys) throw Error(E_INVAL, "null argument");
        if (!(kbuf > 0) || !std::isfinite(kbuf)) throw Error(E_INVAL, "kbuf must be positive");
        HIPCHECK(hipSetDevice(sys->b.ctx->device));
        optimize_rns(sys->b, kbuf, rns_out);
    });
}

mdc_status mdc_sys_set_rebuild(mdc_sys* sys, int automatic) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys) throw Error(E_INVAL, "null argument");
        sys->b.manual_rebuild = automatic == 0;
    });
}

mdc_status mdc_sys_set_boundary(mdc_sys* sys, int periodic_boundary) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys) throw Error(E_INVAL, "null argument");
        Batch& b = sys->b;
        HIPCHECK(hipSetDevice(b.ctx->device));
        for (int s = 0; s < b.nsys; ++s) {
            b.periodic_bc[s] = periodic_boundary ? 1 : 0;
            SysDesc& d = b.desc[s];
            d.any_per = 0;
            for (int a = 0; a < 3; ++a) d.any_per |= (d.per[a] && periodic_boundary) ? 1 : 0;
        }
        HIPCHECK(hipMemcpyAsync(b.d_desc, b.desc.data(), sizeof(SysDesc) * b.nsys, hipMemcpyHostToDevice,
                                b.ctx->stream));
        HIPCHECK(hipStreamSynchronize(b.ctx->stream));
    });
}

mdc_status mdc_sys_set_thermostat(mdc_sys* sys, int count, const double* params, uint64_t seed) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys || (count > 0 && !params)) throw Error(E_INVAL, "null argument");
        if (count < 0 || count > kMaxTherm) throw Error(E_INVAL, "between 0 and 4 thermostats");
        ThermArgs T{};
        T.count = count;
        T.seed = seed;
        // thermostat/base.h:10
        const double kb = 0.008314462175;
        for (int k = 0; k < count; ++k) {
            const double* q = params + 11 * k;
            const double beta = q[0], temp = q[1], dt = q[2];
            if (!(dt > 0) || !(temp >= 0) || !std::isfinite(beta)) throw Error(E_INVAL, "bad Langevin parameters");
            LangevinP& L = T.t[k];
            L.beta = beta;
            // langevin.cxx:25
            L.sigma = std::sqrt(2.0 * beta * kb * temp / dt);
            for (int c = 0; c < 3; ++c) {
                L.heat[c] = q[3 + c];
                L.U[c] = q[6 + c];
            }
            L.xmin = q[9];
            L.xmax = q[10];
        }
        Batch& b = sys->b;
        if (seed != b.therm.seed) b.therm_calls = 0;
        b.therm = T;
    });
}

mdc_status mdc_stage(mdc_sys* sys, const mdc_pot* pot, int stage, int precision) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys || !pot) throw Error(E_INVAL, "null argument");
        if (precision != MDC_FP64 && precision != MDC_MIXED) throw Error(E_INVAL, "precision must be FP64 or MIXED");
        HIPCHECK(hipSetDevice(sys->b.ctx->device));
        check_kinds(sys->b, pot->p);
        stage_compute(sys->b, pot->p, stage, precision);
    });
}

mdc_status mdc_nlist_export(mdc_sys* sys, uint64_t* offsets, uint64_t* ids, uint64_t* total) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys) throw Error(E_INVAL, "null argument");
        Batch& b = sys->b;
        if (!b.list_valid) throw Error(E_STATE, "no neighbour list: call mdc_nlist_build first");
        HIPCHECK(hipSetDevice(b.ctx->device));
        const int n = b.n;
        std::vector<int> nn(n), orig(n), where(n);
        std::vector<int> nl(b.nl_elems);
        hipStream_t st = b.ctx->stream;
        HIPCHECK(hipMemcpyAsync(nn.data(), b.nn, sizeof(int) * n, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(orig.data(), b.orig[b.sc], sizeof(int) * n, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(where.data(), b.where, sizeof(int) * n, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(nl.data(), b.nl, sizeof(int) * b.nl_elems, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        uint64_t tot = 0;
        for (int i = 0; i < n; ++i) tot += (uint64_t)std::min(nn[i], b.cap) + 1;
        if (total) *total = tot;
        if (!offsets || !ids) return;
        uint64_t o = 0;
        for (int a = 0; a < n; ++a) {
            offsets[a] = o;
            const int p = where[a];
            ids[o++] = (uint64_t)a;  // the reference's list includes self
            const int c = std::min(nn[p], b.cap);
            for (int k = 0; k < c; ++k) ids[o++] = (uint64_t)orig[ent_j(nl[list_index(p, k, b.tpa, b.M)])];
        }
        offsets[n] = o;
    });
}

mdc_status mdc_nlist_stats(mdc_sys* sys, double* mean, int* maxn, int* capacity, int* rebuilds) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys) throw Error(E_INVAL, "null argument");
        Batch& b = sys->b;
        HIPCHECK(hipSetDevice(b.ctx->device));
        std::vector<int> nn(b.n);
        HIPCHECK(hipMemcpyAsync(nn.data(), b.nn, sizeof(int) * b.n, hipMemcpyDeviceToHost, b.ctx->stream));
        HIPCHECK(hipStreamSynchronize(b.ctx->stream));
        double s = 0;
        int mx = 0;
        for (int x : nn) {
            s += x;
            mx = std::max(mx, x);
        }
        // per live record (void records have no list)
        if (mean) *mean = b.live ? s / b.live : 0.0;
        if (maxn) *maxn = mx;
        if (capacity) *capacity = b.cap;
        if (rebuilds) *rebuilds = b.rebuilds;
    });
}

mdc_status mdc_compute(mdc_sys* sys, const mdc_pot* pot, uint32_t what, int precision) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys || !pot) throw Error(E_INVAL, "null argument");
        if (precision != MDC_FP64 && precision != MDC_MIXED) throw Error(E_INVAL, "precision must be FP64 or MIXED");
        HIPCHECK(hipSetDevice(sys->b.ctx->device));
        check_kinds(sys->b, pot->p);
        force_compute(sys->b, pot->p, (int)what, precision);
    });
}

mdc_status mdc_compute_noalloc(mdc_sys* sys, const mdc_pot* pot, uint32_t what, int precision) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys || !pot) throw Error(E_INVAL, "null argument");
        if (precision != MDC_FP64 && precision != MDC_MIXED) throw Error(E_INVAL, "precision must be FP64 or MIXED");
        Batch& b = sys->b;
        if (pot->p.type != PT_BPNN)
            throw Error(E_INVAL, "mdc_compute_noalloc evaluates BPNN potentials (the only ones with work arrays)");
        if (!b.list_valid) throw Error(E_STATE, "no neighbour list: build it first (this call builds none)");
        char* base = scratch_on(b.ctx->device);
        if (!base) throw Error(E_STATE, "no work area on this device");
        const size_t need = bpnn_fixed_bytes(b, pot->p, (int)what, precision);
        if (need > kNoAllocBytes)
            throw Error(E_INVAL, "evaluation needs " + std::to_string((need + (1 << 20) - 1) >> 20) +
                                     " MiB of work arrays; the allocation-free work area holds " +
                                     std::to_string(kNoAllocBytes >> 20) + " MiB");
        HIPCHECK(hipSetDevice(b.ctx->device));
        check_kinds(b, pot->p);
        // timing events are created on demand: off for this call
        const bool timing = b.timing;
        b.timing = false;
        b.bp_fixed = base;
        try {
            force_compute(b, pot->p, (int)what, precision);
        } catch (...) {
            b.bp_fixed = nullptr;
            b.timing = timing;
            throw;
        }
        b.bp_fixed = nullptr;
        b.timing = timing;
    });
}

mdc_status mdc_sys_autotune(mdc_sys* sys, const mdc_pot* pot, int precision, int reps, int* best_lanes,
                            double* us_per_eval) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys || !pot || !best_lanes) throw Error(E_INVAL, "null argument");
        if (precision != MDC_FP64 && precision != MDC_MIXED) throw Error(E_INVAL, "precision must be FP64 or MIXED");
        if (reps < 1) throw Error(E_INVAL, "reps must be >= 1");
        Batch& b = sys->b;
        HIPCHECK(hipSetDevice(b.ctx->device));
        check_kinds(b, pot->p);
        const int cand[5] = {1, 2, 4, 8, 16};
        hipStream_t st = b.ctx->stream;
        hipEvent_t e0, e1;
        HIPCHECK(hipEventCreate(&e0));
        HIPCHECK(hipEventCreate(&e1));
        double best = 1e300;
        int best_l = b.tpa;
        const bool lds0 = b.q_lds;
        bool best_lds = lds0, lds_l = lds0;
        try {
            for (int c = 0; c < 5; ++c) {
                const int l = cand[c];
                set_lanes(b, l);
                nlist_rebuild(b);
                force_compute(b, pot->p, MDC_FORCE, precision);  // warm-up, and the force to start from
                double us = 0.0;
                if (pot->p.type == PT_BPNN) {
                    HIPCHECK(hipEventRecord(e0, st));
                    for (int r = 0; r < reps; ++r) force_compute(b, pot->p, MDC_FORCE, precision);
                    HIPCHECK(hipEventRecord(e1, st));
                    HIPCHECK(hipEventSynchronize(e1));
                    float ms = 0.f;
                    HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
                    us = 1e3 * ms / reps;
                } else {
                    // pair potentials: the kernels a step runs (density and
                    // force in step mode), back to back, without host syncs;
                    // on the fixed-point path both table placements
                    double t[2];
                    b.q_lds = false;
                    time_step_kernels(b, pot->p, precision, 1e-6, std::max(reps, 20), t);
                    us = t[0] + t[1];
                    lds_l = false;
                    if (b.q_last) {
                        b.q_lds = true;
                        time_step_kernels(b, pot->p, precision, 1e-6, std::max(reps, 20), t);
                        if (t[0] + t[1] < us) {
                            us = t[0] + t[1];
                            lds_l = true;
                        }
                    }
                }
                if (us_per_eval) us_per_eval[c] = us;
                if (us < best) {
                    best = us;
                    best_l = l;
                    best_lds = lds_l;
                }
            }
        } catch (...) {
            b.q_lds = lds0;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        HIPCHECK(hipEventDestroy(e0));
        HIPCHECK(hipEventDestroy(e1));
        b.q_lds = best_lds;
        set_lanes(b, best_l);
        nlist_rebuild(b);
        force_compute(b, pot->p, MDC_FORCE, precision);
        *best_lanes = best_l;
    });
}

mdc_status mdc_sys_set_block(mdc_sys* sys, int threads) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys) throw Error(E_INVAL, "null argument");
        if (threads != 64 && threads != 128 && threads != 256)
            throw Error(E_INVAL, "threads per workgroup must be 64, 128 or 256");
        sys->b.block = threads;
    });
}

mdc_status mdc_sys_step_path(mdc_sys* sys, int set, int* last) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys) throw Error(E_INVAL, "null argument");
        if (set < -1 || set > 2)
            throw Error(E_INVAL, "set must be -1 (keep), 0 (fixed point, global tables), 1 (FP64 differences) or "
                                 "2 (fixed point, tables in LDS)");
        if (set >= 0) {
            sys->b.q_off = set == 1;
            if (set != 1) sys->b.q_lds = set == 2;
        }
        if (last) *last = sys->b.q_last;
    });
}

mdc_status mdc_sys_autotune_tiles(mdc_sys* sys, const mdc_pot* pot, int precision, int reps, int* best_lanes,
                                  int* best_threads, double* us_out) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys || !pot || !best_lanes || !best_threads) throw Error(E_INVAL, "null argument");
        if (precision != MDC_FP64 && precision != MDC_MIXED) throw Error(E_INVAL, "precision must be FP64 or MIXED");
        if (reps < 1) throw Error(E_INVAL, "reps must be >= 1");
        if (pot->p.type == PT_BPNN) throw Error(E_INVAL, "the tile autotuner covers the LJ and EAM kernels");
        Batch& b = sys->b;
        HIPCHECK(hipSetDevice(b.ctx->device));
        check_kinds(b, pot->p);
        const int lanes[5] = {1, 2, 4, 8, 16}, threads[3] = {64, 128, 256};
        double best = 1e300;
        int bl = b.tpa, bt = b.block;
        for (int c = 0; c < 5; ++c) {
            set_lanes(b, lanes[c]);
            nlist_rebuild(b);
            force_compute(b, pot->p, MDC_FORCE, precision);  // warm-up, and the force to start from
            for (int q = 0; q < 3; ++q) {
                b.block = threads[q];
                double t[2];
                time_step_kernels(b, pot->p, precision, 1e-6, std::max(reps, 20), t);
                const double us = t[0] + t[1];
                if (us_out) us_out[3 * c + q] = us;
                if (us < best) {
                    best = us;
                    bl = lanes[c];
                    bt = threads[q];
                }
            }
        }
        set_lanes(b, bl);
        b.block = bt;
        nlist_rebuild(b);
        force_compute(b, pot->p, MDC_FORCE, precision);
        *best_lanes = bl;
        *best_threads = bt;
    });
}

mdc_status mdc_sys_set_graphs(mdc_sys* sys, int on) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys) throw Error(E_INVAL, "null argument");
        HIPCHECK(hipSetDevice(sys->b.ctx->device));
        sys->b.graphs = on != 0;
        if (!sys->b.graphs) graph_release(sys->b);
    });
}

mdc_status mdc_sys_graph_stats(const mdc_sys* sys, uint64_t* captures, uint64_t* launches) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys) throw Error(E_INVAL, "null argument");
        if (captures) *captures = (uint64_t)sys->b.graph_captures;
        if (launches) *launches = (uint64_t)sys->b.graph_launches;
    });
}

mdc_status mdc_sys_time_kernels(mdc_sys* sys, const mdc_pot* pot, int precision, double dt, int reps,
                                double* us_out) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys || !pot || !us_out) throw Error(E_INVAL, "null argument");
        if (precision != MDC_FP64 && precision != MDC_MIXED) throw Error(E_INVAL, "precision must be FP64 or MIXED");
        if (!(dt > 0) || reps < 1) throw Error(E_INVAL, "dt must be positive and reps >= 1");
        HIPCHECK(hipSetDevice(sys->b.ctx->device));
        check_kinds(sys->b, pot->p);
        time_step_kernels(sys->b, pot->p, precision, dt, reps, us_out);
    });
}

mdc_status mdc_step(mdc_sys* sys, const mdc_pot* pot, double dt, int nsteps, int precision, int* rebuilds) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys || !pot) throw Error(E_INVAL, "null argument");
        if (precision != MDC_FP64 && precision != MDC_MIXED) throw Error(E_INVAL, "precision must be FP64 or MIXED");
        if (!(dt > 0) || nsteps < 0) throw Error(E_INVAL, "dt must be positive and nsteps >= 0");
        HIPCHECK(hipSetDevice(sys->b.ctx->device));
        check_kinds(sys->b, pot->p);
        int r0 = sys->b.rebuilds;
        run_steps(sys->b, pot->p, dt, nsteps, precision, 16);
        if (rebuilds) *rebuilds = sys->b.rebuilds - r0;
    });
}

mdc_status mdc_thermo(mdc_sys* sys, double* out) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys || !out) throw Error(E_INVAL, "null argument");
        HIPCHECK(hipSetDevice(sys->b.ctx->device));
        thermo(sys->b, out);
    });
}

mdc_status mdc_sync(mdc_sys* sys) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys) throw Error(E_INVAL, "null argument");
        HIPCHECK(hipStreamSynchronize(sys->b.ctx->stream));
    });
}

mdc_status mdc_bpnn_inspect(mdc_sys* sys, const mdc_pot* pot, double* G, double* dEdG, int* nin) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys || !pot || pot->p.type != PT_BPNN) throw Error(E_INVAL, "needs a system and a BPNN potential");
        if (nin) *nin = bpnn_inputs(pot->p);
        if (G || dEdG) bpnn_inspect(sys->b, pot->p, G, dEdG);
    });
}

mdc_status mdc_copy_calibration(int device, uint64_t bytes, int reps, double* us_per_copy) {
    return guard(nullptr, [&] { copy_calibration(device, (size_t)bytes, reps, us_per_copy); });
}

mdc_status mdc_timing(mdc_sys* sys, int enable, double* total_ms, int64_t* launches) {
    return guard(sys ? sys->b.ctx : nullptr, [&] {
        if (!sys) throw Error(E_INVAL, "null argument");
        Batch& b = sys->b;
        HIPCHECK(hipStreamSynchronize(b.ctx->stream));
        timer_drain(b);
        for (int k = 0; k < kKinds; ++k) {
            if (total_ms) total_ms[k] = b.kt_ms[k];
            if (launches) launches[k] = b.kt_n[k];
        }
        if (enable >= 0) {
            b.timing = enable != 0;
            for (int k = 0; k < kKinds; ++k) {
                b.kt_ms[k] = 0;
                b.kt_n[k] = 0;
            }
        }
    });
}

}  // extern "C"
