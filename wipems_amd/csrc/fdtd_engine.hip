// fdtd_engine.hip -- MI355X (gfx950) FDTD engine behind the C ABI of
// include/wipems_fdtd.h.
//
// One handle = one Cartesian EC-FDTD domain on one GPU.  Fields live in HBM as
// SoA arrays [3][Nx][Ny][Nzp] (z contiguous, padded to Nzp), so a wave of 64
// lanes reads 64 consecutive z values of one component in one coalesced
// stream.  A timestep is a short chain of kernels on the handle's stream:
//
//   ade_pre_volt  (sparse)  Engine_Ext_LorentzMaterial::DoPreVoltageUpdates
//   volt update   (dense)   Engine::UpdateVoltages with UPML Pre/Post fused
//   tfsf volt     (sparse)  Engine_Ext_TFSF::DoPostVoltageUpdates
//   ade_apply_volt(sparse)  Engine_Ext_Dispersive::Apply2Voltages
//   exc volt      (sparse)  Engine_Ext_Excitation::Apply2Voltages
//   ade_pre_curr / curr update / tfsf curr / ade_apply_curr / exc curr
//   step_end      (tiny)    ++numTS, probe integrals, FD accumulation (FP64)
//
// This is the order of Engine::IterateTS (FDTD/engine.cpp:255-274) with the
// extensions sorted by priority (FDTD/engine.cpp:85-96,212-253): pre hooks run
// Excitation, Lorentz, TFSF, UPML; post/apply hooks run UPML, TFSF, Lorentz,
// Excitation.  The timestep counter lives in device memory so the whole step
// can be captured once in a HIP graph and replayed.
//
// Arithmetic follows the reference operation order exactly,
//   V *= vv;  V += vi * (((a - b) - c) + d)         (FDTD/engine.cpp:126-134)
// with contraction disabled (no FMA) and, for FP32, denormals flushed (the
// SSE engines set FTZ|DAZ, tools/denormal.h:19-31), so an FP32 handle
// reproduces the reference FP32 engine bit for bit.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>

#include "wipems_fdtd.h"

#pragma clang fp contract(off)

#define FDTD_VERSION_STR "wipems-fdtd 0.1 (gfx950)"

static const double kEPS0 = 8.85418781762e-12;   // tools/constants.h:22
static const double kMUE0 = 1.256637062e-6;      // tools/constants.h:23

// ---------------------------------------------------------------------------
// error handling
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;

static int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIPCHK(expr)                                                                 \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess)                                                        \
            return fail(FDTD_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------------------
// storage types: FP32/FP64 store the arithmetic type; FP16/BF16 store a
// reduced type and compute in float with per-component scales.
// ---------------------------------------------------------------------------
template <typename S> struct StoreTraits;
template <> struct StoreTraits<float> {
    typedef float A;
    __device__ static inline float load(const float* p, float) { return *p; }
    __device__ static inline void store(float* p, float v, float) { *p = v; }
};
template <> struct StoreTraits<double> {
    typedef double A;
    __device__ static inline double load(const double* p, double) { return *p; }
    __device__ static inline void store(double* p, double v, double) { *p = v; }
};
template <> struct StoreTraits<__half> {
    typedef float A;
    __device__ static inline float load(const __half* p, float s) { return __half2float(*p) * s; }
    __device__ static inline void store(__half* p, float v, float inv_s) { *p = __float2half_rn(v * inv_s); }
};
template <> struct StoreTraits<__hip_bfloat16> {
    typedef float A;
    __device__ static inline float load(const __hip_bfloat16* p, float s) { return __bfloat162float(*p) * s; }
    __device__ static inline void store(__hip_bfloat16* p, float v, float inv_s) { *p = __float2bfloat16(v * inv_s); }
};

// ---------------------------------------------------------------------------
// kernel argument blocks
// ---------------------------------------------------------------------------
struct Dims {
    int nx, ny, nz;        // mesh lines
    long long sx, sy;      // element strides of x and y (z stride is 1)
    long long sc;          // component stride
};

template <typename A>
struct Scales {            // per-component storage scales (1 for FP32/FP64)
    A v[3], iv[3];         // scale and its inverse, voltage
    A c[3], ic[3];         // scale and its inverse, current
};

#define MAX_PML_SLABS 6
// A slab's arrays are stored [3][nx][ny][nzp] with the z range widened to
// start at za = s[2] rounded down to a multiple of 16 and padded to nzp, so
// a 16-byte-aligned vector of consecutive z cells of the main arrays maps to
// an aligned vector of the slab arrays.  Padding cells hold zeros.
struct PmlSlab {
    int s[3];              // first line per axis
    int n[3];              // lines per axis
    int za;                // aligned first z of the slab arrays
    int nzp;               // padded z extent of the slab arrays
    long long off;         // element offset of this slab's component 0
    long long sc;          // component stride inside the slab (nx*ny*nzp)
};
struct PmlSet {
    int count;
    PmlSlab slab[MAX_PML_SLABS];
};

template <typename A>
struct PmlPtrs {
    const A* vv; const A* vvfo; const A* vvfn;   // voltage side
    const A* ii; const A* iifo; const A* iifn;   // current side
    A* vflux; A* cflux;
};

// coefficient access modes
enum { CF_FULL = 0, CF_IDX16 = 1, CF_IDX32 = 2 };

template <typename A>
struct CoeffPtrs {
    const A* a;            // full: vv (E) or ii (H) [3][cells]; table: [U][8] rows {a0,a1,a2,_,b0,b1,b2,_}
    const A* b;            // full: vi (E) or iv (H)
    const void* idx;       // per-cell index (compressed modes)
};

__device__ __forceinline__ int find_slab(const PmlSet& ps, int x, int y, int z, long long& local) {
    for (int s = 0; s < ps.count; ++s) {
        const PmlSlab& b = ps.slab[s];
        int lx = x - b.s[0], ly = y - b.s[1], lz = z - b.s[2];
        if ((unsigned)lx < (unsigned)b.n[0] && (unsigned)ly < (unsigned)b.n[1] && (unsigned)lz < (unsigned)b.n[2]) {
            local = b.off + ((long long)lx * b.n[1] + ly) * b.nzp + (z - b.za);
            return s;
        }
    }
    return -1;
}

// slab holding any of the cells z0..z0+vz-1 at (x,y); base = slab-array
// offset of z0 (aligned like z0), mask bit k = cell z0+k lies in the slab
__device__ __forceinline__ int find_slab_vec(const PmlSet& ps, int x, int y, int z0, int vz, long long& base, unsigned& mask) {
    for (int s = 0; s < ps.count; ++s) {
        const PmlSlab& b = ps.slab[s];
        const int lx = x - b.s[0], ly = y - b.s[1];
        if ((unsigned)lx >= (unsigned)b.n[0] || (unsigned)ly >= (unsigned)b.n[1]) continue;
        unsigned m = 0;
        for (int k = 0; k < vz; ++k)
            if ((unsigned)(z0 + k - b.s[2]) < (unsigned)b.n[2]) m |= 1u << k;
        if (!m) continue;
        base = b.off + ((long long)lx * b.n[1] + ly) * b.nzp + (z0 - b.za);
        mask = m;
        return s;
    }
    return -1;
}

template <typename A, int CM>
__device__ __forceinline__ void load_coeffs(const CoeffPtrs<A>& cf, long long p, long long sc, A a[3], A b[3]) {
    if (CM == CF_FULL) {
#pragma unroll
        for (int n = 0; n < 3; ++n) { a[n] = cf.a[n * sc + p]; b[n] = cf.b[n * sc + p]; }
    } else {
        unsigned c = (CM == CF_IDX16) ? (unsigned)((const unsigned short*)cf.idx)[p]
                                      : ((const unsigned*)cf.idx)[p];
        const A* row = cf.a + (size_t)c * 8;
#pragma unroll
        for (int n = 0; n < 3; ++n) { a[n] = row[n]; b[n] = row[4 + n]; }
    }
}

// ---------------------------------------------------------------------------
// main voltage update (Engine::UpdateVoltages, FDTD/engine.cpp:108-161, with the
// z-boundary of Engine_SSE_Compressed, FDTD/engine_sse_compressed.cpp:95-153:
// the z-1 neighbour at z=0 is a zero-filled lane) with UPML Pre/Post fused
// (FDTD/extensions/engine_ext_upml.cpp:57-135):
//   pre : h = pv*V - fo*F ; V = F ; F = h
//   main: V = V*vv + vi*curl
//   post: h = F ; F = V ; V = h + fn*F
// ---------------------------------------------------------------------------
template <typename S, int CM, bool PML>
__device__ __forceinline__ void volt_cell(S* __restrict__ V, const S* __restrict__ I, const CoeffPtrs<typename StoreTraits<S>::A>& cf,
                                          const Dims& d, const Scales<typename StoreTraits<S>::A>& sc, const PmlSet& ps,
                                          const PmlPtrs<typename StoreTraits<S>::A>& pp, int x, int y, int z) {
    typedef StoreTraits<S> T;
    typedef typename T::A A;
    const long long p = x * d.sx + y * d.sy + z;
    const long long c1 = d.sc, c2 = 2 * d.sc;
    const long long ox = x > 0 ? d.sx : 0;   // x-1 neighbour, self at x=0 (shift[0])
    const long long oy = y > 0 ? d.sy : 0;   // y-1 neighbour, self at y=0 (shift[1])

    const A i0 = T::load(I + p, sc.c[0]);
    const A i1 = T::load(I + c1 + p, sc.c[1]);
    const A i2 = T::load(I + c2 + p, sc.c[2]);
    const A i2_ym = T::load(I + c2 + p - oy, sc.c[2]);
    const A i0_ym = T::load(I + p - oy, sc.c[0]);
    const A i2_xm = T::load(I + c2 + p - ox, sc.c[2]);
    const A i1_xm = T::load(I + c1 + p - ox, sc.c[1]);
    const A i1_zm = z > 0 ? T::load(I + c1 + p - 1, sc.c[1]) : A(0);
    const A i0_zm = z > 0 ? T::load(I + p - 1, sc.c[0]) : A(0);

    A curl[3];
    curl[0] = ((i2 - i2_ym) - i1) + i1_zm;
    curl[1] = ((i0 - i0_zm) - i2) + i2_xm;
    curl[2] = ((i1 - i1_xm) - i0) + i0_ym;

    A vv[3], vi[3];
    load_coeffs<A, CM>(cf, p, d.sc, vv, vi);

    long long lp = 0;
    int slab = PML ? find_slab(ps, x, y, z, lp) : -1;
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        S* vp = V + n * d.sc + p;
        A v = T::load(vp, sc.v[n]);
        if (PML && slab >= 0) {
            const PmlSlab& b = ps.slab[slab];
            const long long q = lp + n * b.sc;
            A F = pp.vflux[q];
            A h = pp.vv[q] * v - pp.vvfo[q] * F;
            A t = F;
            t = t * vv[n];
            t = t + vi[n] * curl[n];
            pp.vflux[q] = t;
            v = h + pp.vvfn[q] * t;
        } else {
            v = v * vv[n];
            v = v + vi[n] * curl[n];
        }
        T::store(vp, v, sc.iv[n]);
    }
}

template <typename S, int CM, bool PML>
__global__ void __launch_bounds__(256)
k_volt(S* __restrict__ V, const S* __restrict__ I, CoeffPtrs<typename StoreTraits<S>::A> cf, Dims d,
       Scales<typename StoreTraits<S>::A> sc, PmlSet ps, PmlPtrs<typename StoreTraits<S>::A> pp) {
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int x = blockIdx.z;
    if (z >= d.nz || y >= d.ny) return;
    volt_cell<S, CM, PML>(V, I, cf, d, sc, ps, pp, x, y, z);
}

// ---------------------------------------------------------------------------
// main current update (Engine::UpdateCurrents, FDTD/engine.cpp:163-210) over
// x<Nx-1, y<Ny-1, z<Nz-1 with UPML fused (engine_ext_upml.cpp:141-234).
// ---------------------------------------------------------------------------
template <typename S, int CM, bool PML>
__device__ __forceinline__ void curr_cell(S* __restrict__ I, const S* __restrict__ V, const CoeffPtrs<typename StoreTraits<S>::A>& cf,
                                          const Dims& d, const Scales<typename StoreTraits<S>::A>& sc, const PmlSet& ps,
                                          const PmlPtrs<typename StoreTraits<S>::A>& pp, int x, int y, int z) {
    typedef StoreTraits<S> T;
    typedef typename T::A A;
    const long long p = x * d.sx + y * d.sy + z;
    const long long c1 = d.sc, c2 = 2 * d.sc;

    const A v0 = T::load(V + p, sc.v[0]);
    const A v1 = T::load(V + c1 + p, sc.v[1]);
    const A v2 = T::load(V + c2 + p, sc.v[2]);
    const A v2_yp = T::load(V + c2 + p + d.sy, sc.v[2]);
    const A v0_yp = T::load(V + p + d.sy, sc.v[0]);
    const A v2_xp = T::load(V + c2 + p + d.sx, sc.v[2]);
    const A v1_xp = T::load(V + c1 + p + d.sx, sc.v[1]);
    const A v1_zp = T::load(V + c1 + p + 1, sc.v[1]);
    const A v0_zp = T::load(V + p + 1, sc.v[0]);

    A curl[3];
    curl[0] = ((v2 - v2_yp) - v1) + v1_zp;
    curl[1] = ((v0 - v0_zp) - v2) + v2_xp;
    curl[2] = ((v1 - v1_xp) - v0) + v0_yp;

    A ii[3], iv[3];
    load_coeffs<A, CM>(cf, p, d.sc, ii, iv);

    long long lp = 0;
    int slab = PML ? find_slab(ps, x, y, z, lp) : -1;
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        S* ip = I + n * d.sc + p;
        A c = T::load(ip, sc.c[n]);
        if (PML && slab >= 0) {
            const PmlSlab& b = ps.slab[slab];
            const long long q = lp + n * b.sc;
            A F = pp.cflux[q];
            A h = pp.ii[q] * c - pp.iifo[q] * F;
            A t = F;
            t = t * ii[n];
            t = t + iv[n] * curl[n];
            pp.cflux[q] = t;
            c = h + pp.iifn[q] * t;
        } else {
            c = c * ii[n];
            c = c + iv[n] * curl[n];
        }
        T::store(ip, c, sc.ic[n]);
    }
}

template <typename S, int CM, bool PML>
__global__ void __launch_bounds__(256)
k_curr(S* __restrict__ I, const S* __restrict__ V, CoeffPtrs<typename StoreTraits<S>::A> cf, Dims d,
       Scales<typename StoreTraits<S>::A> sc, PmlSet ps, PmlPtrs<typename StoreTraits<S>::A> pp) {
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int x = blockIdx.z;
    if (z >= d.nz - 1 || y >= d.ny - 1 || x >= d.nx - 1) return;
    curr_cell<S, CM, PML>(I, V, cf, d, sc, ps, pp, x, y, z);
}

// ---------------------------------------------------------------------------
// vectorised updates: each lane updates VZ consecutive z cells with one wide
// load/store per field component (16 bytes for FP32, FP16/BF16 and FP64), so
// the number of memory instructions per cell drops VZ-fold; the z-1 (E) and
// z+1 (H) neighbours come from the neighbouring element in the same vector,
// and only the first/last element needs one scalar load.  Per-cell arithmetic
// is the same as volt_cell/curr_cell.
// ---------------------------------------------------------------------------
template <typename S, int VZ>
struct alignas(sizeof(S) * VZ) SVec {
    S v[VZ];
};

template <typename S, int VZ>
__device__ __forceinline__ void vload(const S* p, typename StoreTraits<S>::A* out, typename StoreTraits<S>::A s) {
    const SVec<S, VZ> t = *reinterpret_cast<const SVec<S, VZ>*>(p);
#pragma unroll
    for (int k = 0; k < VZ; ++k) out[k] = StoreTraits<S>::load(&t.v[k], s);
}

template <typename S, int VZ>
__device__ __forceinline__ void vstore(S* p, const typename StoreTraits<S>::A* in, typename StoreTraits<S>::A inv_s, int valid) {
    SVec<S, VZ> t;
#pragma unroll
    for (int k = 0; k < VZ; ++k) StoreTraits<S>::store(&t.v[k], in[k], inv_s);
    if (valid >= VZ) {
        *reinterpret_cast<SVec<S, VZ>*>(p) = t;
    } else {
        for (int k = 0; k < valid; ++k) p[k] = t.v[k];
    }
}

template <typename A, int CM, int VZ>
__device__ __forceinline__ void vload_coeffs(const CoeffPtrs<A>& cf, long long p, long long sc, A a[3][VZ], A b[3][VZ]) {
    if (CM == CF_FULL) {
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            const SVec<A, VZ> ta = *reinterpret_cast<const SVec<A, VZ>*>(cf.a + n * sc + p);
            const SVec<A, VZ> tb = *reinterpret_cast<const SVec<A, VZ>*>(cf.b + n * sc + p);
#pragma unroll
            for (int k = 0; k < VZ; ++k) { a[n][k] = ta.v[k]; b[n][k] = tb.v[k]; }
        }
    } else {
        unsigned c[VZ];
        if (CM == CF_IDX16) {
            const SVec<unsigned short, VZ> ti = *reinterpret_cast<const SVec<unsigned short, VZ>*>((const unsigned short*)cf.idx + p);
#pragma unroll
            for (int k = 0; k < VZ; ++k) c[k] = ti.v[k];
        } else {
            const SVec<unsigned, VZ> ti = *reinterpret_cast<const SVec<unsigned, VZ>*>((const unsigned*)cf.idx + p);
#pragma unroll
            for (int k = 0; k < VZ; ++k) c[k] = ti.v[k];
        }
#pragma unroll
        for (int k = 0; k < VZ; ++k) {
            const A* row = cf.a + (size_t)c[k] * 8;
            const SVec<A, 4> r0 = *reinterpret_cast<const SVec<A, 4>*>(row);
            const SVec<A, 4> r1 = *reinterpret_cast<const SVec<A, 4>*>(row + 4);
#pragma unroll
            for (int n = 0; n < 3; ++n) { a[n][k] = r0.v[n]; b[n][k] = r1.v[n]; }
        }
    }
}

template <typename S, int CM, bool PML, int VZ>
__global__ void __launch_bounds__(256)
k_volt_vec(S* __restrict__ V, const S* __restrict__ I, CoeffPtrs<typename StoreTraits<S>::A> cf, Dims d,
           Scales<typename StoreTraits<S>::A> sc, PmlSet ps, PmlPtrs<typename StoreTraits<S>::A> pp) {
    typedef StoreTraits<S> T;
    typedef typename T::A A;
    // flat (y, z) index over the padded rows of one x plane
    const long long q0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * VZ;
    const int y = (int)(q0 / d.sy);
    const int z0 = (int)(q0 - (long long)y * d.sy);
    const int x = blockIdx.y;
    if (y >= d.ny || z0 >= d.nz) return;
    const long long p = x * d.sx + q0;
    const long long c1 = d.sc, c2 = 2 * d.sc;
    const long long ox = x > 0 ? d.sx : 0;
    const long long oy = y > 0 ? d.sy : 0;
    A i0[VZ], i1[VZ], i2[VZ], i2_ym[VZ], i0_ym[VZ], i2_xm[VZ], i1_xm[VZ];
    vload<S, VZ>(I + p, i0, sc.c[0]);
    vload<S, VZ>(I + c1 + p, i1, sc.c[1]);
    vload<S, VZ>(I + c2 + p, i2, sc.c[2]);
    vload<S, VZ>(I + c2 + p - oy, i2_ym, sc.c[2]);
    vload<S, VZ>(I + p - oy, i0_ym, sc.c[0]);
    vload<S, VZ>(I + c2 + p - ox, i2_xm, sc.c[2]);
    vload<S, VZ>(I + c1 + p - ox, i1_xm, sc.c[1]);
    const A i1_zm0 = z0 > 0 ? T::load(I + c1 + p - 1, sc.c[1]) : A(0);
    const A i0_zm0 = z0 > 0 ? T::load(I + p - 1, sc.c[0]) : A(0);
    A curl[3][VZ];
#pragma unroll
    for (int k = 0; k < VZ; ++k) {
        const A i1_zm = k == 0 ? i1_zm0 : i1[k - 1];
        const A i0_zm = k == 0 ? i0_zm0 : i0[k - 1];
        curl[0][k] = ((i2[k] - i2_ym[k]) - i1[k]) + i1_zm;
        curl[1][k] = ((i0[k] - i0_zm) - i2[k]) + i2_xm[k];
        curl[2][k] = ((i1[k] - i1_xm[k]) - i0[k]) + i0_ym[k];
    }
    A vv[3][VZ], vi[3][VZ];
    vload_coeffs<A, CM, VZ>(cf, p, d.sc, vv, vi);
    const int valid = min(VZ, d.nz - z0);
    long long pb = 0;
    unsigned pm = 0;
    const int slab = PML ? find_slab_vec(ps, x, y, z0, VZ, pb, pm) : -1;
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        A v[VZ];
        vload<S, VZ>(V + n * d.sc + p, v, sc.v[n]);
        if (PML && slab >= 0) {
            const long long q = pb + n * ps.slab[slab].sc;
            A F[VZ], kv[VZ], kfo[VZ], kfn[VZ];
            vload<A, VZ>(pp.vflux + q, F, A(1));
            vload<A, VZ>(pp.vv + q, kv, A(1));
            vload<A, VZ>(pp.vvfo + q, kfo, A(1));
            vload<A, VZ>(pp.vvfn + q, kfn, A(1));
#pragma unroll
            for (int k = 0; k < VZ; ++k) {
                if (pm & (1u << k)) {
                    A h = kv[k] * v[k] - kfo[k] * F[k];
                    A t = F[k];
                    t = t * vv[n][k];
                    t = t + vi[n][k] * curl[n][k];
                    F[k] = t;
                    v[k] = h + kfn[k] * t;
                } else {
                    v[k] = v[k] * vv[n][k];
                    v[k] = v[k] + vi[n][k] * curl[n][k];
                }
            }
            vstore<A, VZ>(pp.vflux + q, F, A(1), VZ);
        } else {
#pragma unroll
            for (int k = 0; k < VZ; ++k) {
                v[k] = v[k] * vv[n][k];
                v[k] = v[k] + vi[n][k] * curl[n][k];
            }
        }
        vstore<S, VZ>(V + n * d.sc + p, v, sc.iv[n], valid);
    }
}

template <typename S, int CM, bool PML, int VZ>
__global__ void __launch_bounds__(256)
k_curr_vec(S* __restrict__ I, const S* __restrict__ V, CoeffPtrs<typename StoreTraits<S>::A> cf, Dims d,
           Scales<typename StoreTraits<S>::A> sc, PmlSet ps, PmlPtrs<typename StoreTraits<S>::A> pp) {
    typedef StoreTraits<S> T;
    typedef typename T::A A;
    const long long q0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * VZ;
    const int y = (int)(q0 / d.sy);
    const int z0 = (int)(q0 - (long long)y * d.sy);
    const int x = blockIdx.y;
    if (y >= d.ny - 1 || z0 >= d.nz - 1) return;
    const long long p = x * d.sx + q0;
    const long long c1 = d.sc, c2 = 2 * d.sc;
    A v0[VZ], v1[VZ], v2[VZ], v2_yp[VZ], v0_yp[VZ], v2_xp[VZ], v1_xp[VZ];
    vload<S, VZ>(V + p, v0, sc.v[0]);
    vload<S, VZ>(V + c1 + p, v1, sc.v[1]);
    vload<S, VZ>(V + c2 + p, v2, sc.v[2]);
    vload<S, VZ>(V + c2 + p + d.sy, v2_yp, sc.v[2]);
    vload<S, VZ>(V + p + d.sy, v0_yp, sc.v[0]);
    vload<S, VZ>(V + c2 + p + d.sx, v2_xp, sc.v[2]);
    vload<S, VZ>(V + c1 + p + d.sx, v1_xp, sc.v[1]);
    // z+1 neighbour of the last element: only needed when it is a valid line
    const bool tail = z0 + VZ < d.nz;
    const A v1_zpl = tail ? T::load(V + c1 + p + VZ, sc.v[1]) : A(0);
    const A v0_zpl = tail ? T::load(V + p + VZ, sc.v[0]) : A(0);
    A curl[3][VZ];
#pragma unroll
    for (int k = 0; k < VZ; ++k) {
        const A v1_zp = k == VZ - 1 ? v1_zpl : v1[k + 1];
        const A v0_zp = k == VZ - 1 ? v0_zpl : v0[k + 1];
        curl[0][k] = ((v2[k] - v2_yp[k]) - v1[k]) + v1_zp;
        curl[1][k] = ((v0[k] - v0_zp) - v2[k]) + v2_xp[k];
        curl[2][k] = ((v1[k] - v1_xp[k]) - v0[k]) + v0_yp[k];
    }
    A ii[3][VZ], iv[3][VZ];
    vload_coeffs<A, CM, VZ>(cf, p, d.sc, ii, iv);
    const int valid = min(VZ, d.nz - 1 - z0);
    long long pb = 0;
    unsigned pm = 0;
    const int slab = PML ? find_slab_vec(ps, x, y, z0, VZ, pb, pm) : -1;
    // cells at z >= nz-1 are not updated: keep them out of the slab mask too
    if (valid < VZ) pm &= (1u << valid) - 1u;
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        A c[VZ];
        vload<S, VZ>(I + n * d.sc + p, c, sc.c[n]);
        if (PML && slab >= 0) {
            const long long q = pb + n * ps.slab[slab].sc;
            A F[VZ], ki[VZ], kfo[VZ], kfn[VZ];
            vload<A, VZ>(pp.cflux + q, F, A(1));
            vload<A, VZ>(pp.ii + q, ki, A(1));
            vload<A, VZ>(pp.iifo + q, kfo, A(1));
            vload<A, VZ>(pp.iifn + q, kfn, A(1));
#pragma unroll
            for (int k = 0; k < VZ; ++k) {
                if (pm & (1u << k)) {
                    A h = ki[k] * c[k] - kfo[k] * F[k];
                    A t = F[k];
                    t = t * ii[n][k];
                    t = t + iv[n][k] * curl[n][k];
                    F[k] = t;
                    c[k] = h + kfn[k] * t;
                } else {
                    c[k] = c[k] * ii[n][k];
                    c[k] = c[k] + iv[n][k] * curl[n][k];
                }
            }
            vstore<A, VZ>(pp.cflux + q, F, A(1), VZ);
        } else {
#pragma unroll
            for (int k = 0; k < VZ; ++k) {
                c[k] = c[k] * ii[n][k];
                c[k] = c[k] + iv[n][k] * curl[n][k];
            }
        }
        vstore<S, VZ>(I + n * d.sc + p, c, sc.ic[n], valid);
    }
}

// ---------------------------------------------------------------------------
// banded launch order: one thread per cell like k_volt/k_curr, but block ids
// are remapped so that each XCD (hardware places block b on XCD b % 8) walks
// a contiguous run of logical tiles ordered (y band, x, y tile in band, z
// tile).  A band plane (band_tiles x tile_y rows x Nz) is small, so the x-1
// plane and the y-1 rows a tile needs were just fetched into the same XCD's
// L2 by the tiles before it.
// ---------------------------------------------------------------------------
struct BandGrid {
    int tiles_z, tiles_y, band_tiles, nbands, nx;
    int nblocks, per_xcd;
};

__device__ __forceinline__ bool band_tile(const BandGrid& g, int& tz, int& ty, int& x) {
    const int b = blockIdx.x;
    const int t = (b & 7) * g.per_xcd + (b >> 3);
    if (t >= g.nblocks) return false;
    tz = t % g.tiles_z;
    int r = t / g.tiles_z;
    const int tyb = r % g.band_tiles;
    r /= g.band_tiles;
    x = r % g.nx;
    const int band = r / g.nx;
    ty = band * g.band_tiles + tyb;
    return ty < g.tiles_y;
}

template <typename S, int CM, bool PML>
__global__ void __launch_bounds__(256)
k_volt_band(S* __restrict__ V, const S* __restrict__ I, CoeffPtrs<typename StoreTraits<S>::A> cf, Dims d,
            Scales<typename StoreTraits<S>::A> sc, PmlSet ps, PmlPtrs<typename StoreTraits<S>::A> pp, BandGrid g) {
    int tz, ty, x;
    if (!band_tile(g, tz, ty, x)) return;
    const int z = tz * 64 + threadIdx.x;
    const int y = ty * blockDim.y + threadIdx.y;
    if (z >= d.nz || y >= d.ny) return;
    volt_cell<S, CM, PML>(V, I, cf, d, sc, ps, pp, x, y, z);
}

template <typename S, int CM, bool PML>
__global__ void __launch_bounds__(256)
k_curr_band(S* __restrict__ I, const S* __restrict__ V, CoeffPtrs<typename StoreTraits<S>::A> cf, Dims d,
            Scales<typename StoreTraits<S>::A> sc, PmlSet ps, PmlPtrs<typename StoreTraits<S>::A> pp, BandGrid g) {
    int tz, ty, x;
    if (!band_tile(g, tz, ty, x)) return;
    const int z = tz * 64 + threadIdx.x;
    const int y = ty * blockDim.y + threadIdx.y;
    if (z >= d.nz - 1 || y >= d.ny - 1) return;
    curr_cell<S, CM, PML>(I, V, cf, d, sc, ps, pp, x, y, z);
}

// ---------------------------------------------------------------------------
// x-marching versions of the two updates (2.5-D blocking).  A block is a
// 64(z) x TY(y) tile of columns; each thread walks its (y,z) column through an
// x chunk and keeps the x-neighbour values of the previous plane in
// registers, so every field element is fetched from HBM once per kernel; the
// y and z neighbours are re-read by adjacent lanes/waves of the same block
// and hit L1/L2.  The linear block id is remapped so that consecutive tiles
// (which share their y halo rows) run on the same XCD and share its L2.
// Arithmetic is identical to k_volt / k_curr.
// ---------------------------------------------------------------------------
struct MarchGrid {
    int tiles_z, tiles_y, chunks_x;    // logical tile grid
    int chunk;                         // x planes per chunk
    int nblocks;                       // tiles_z * tiles_y * chunks_x
    int per_xcd;                       // ceil(nblocks / 8)
};

__device__ __forceinline__ bool march_tile(const MarchGrid& g, int& tz, int& ty, int& cx) {
    // hardware dispatch places block b on XCD b % 8; give each XCD a
    // contiguous run of logical tiles (z fastest, then y, then x chunk)
    const int b = blockIdx.x;
    const int t = (b & 7) * g.per_xcd + (b >> 3);
    if (t >= g.nblocks) return false;
    tz = t % g.tiles_z;
    const int r = t / g.tiles_z;
    ty = r % g.tiles_y;
    cx = r / g.tiles_y;
    return true;
}

template <typename S, int CM, bool PML>
__global__ void __launch_bounds__(256)
k_volt_march(S* __restrict__ V, const S* __restrict__ I, CoeffPtrs<typename StoreTraits<S>::A> cf, Dims d,
             Scales<typename StoreTraits<S>::A> sc, PmlSet ps, PmlPtrs<typename StoreTraits<S>::A> pp, MarchGrid g) {
    typedef StoreTraits<S> T;
    typedef typename T::A A;
    int tz, ty, cx;
    if (!march_tile(g, tz, ty, cx)) return;
    const int z = tz * 64 + threadIdx.x;
    const int y = ty * blockDim.y + threadIdx.y;
    if (z >= d.nz || y >= d.ny) return;
    const int x0 = cx * g.chunk;
    const int x1 = min(d.nx, x0 + g.chunk);
    const long long c1 = d.sc, c2 = 2 * d.sc;
    const long long oy = y > 0 ? d.sy : 0;
    long long p = x0 * d.sx + y * d.sy + z;
    // x-1 neighbours of the first plane (the cell itself at x = 0)
    A i1_prev, i2_prev;
    if (x0 > 0) {
        i1_prev = T::load(I + c1 + p - d.sx, sc.c[1]);
        i2_prev = T::load(I + c2 + p - d.sx, sc.c[2]);
    } else {
        i1_prev = T::load(I + c1 + p, sc.c[1]);
        i2_prev = T::load(I + c2 + p, sc.c[2]);
    }
    for (int x = x0; x < x1; ++x, p += d.sx) {
        const A i0 = T::load(I + p, sc.c[0]);
        const A i1 = T::load(I + c1 + p, sc.c[1]);
        const A i2 = T::load(I + c2 + p, sc.c[2]);
        const A i2_ym = T::load(I + c2 + p - oy, sc.c[2]);
        const A i0_ym = T::load(I + p - oy, sc.c[0]);
        const A i1_zm = z > 0 ? T::load(I + c1 + p - 1, sc.c[1]) : A(0);
        const A i0_zm = z > 0 ? T::load(I + p - 1, sc.c[0]) : A(0);
        A curl[3];
        curl[0] = ((i2 - i2_ym) - i1) + i1_zm;
        curl[1] = ((i0 - i0_zm) - i2) + i2_prev;
        curl[2] = ((i1 - i1_prev) - i0) + i0_ym;
        i1_prev = i1;
        i2_prev = i2;

        A vv[3], vi[3];
        load_coeffs<A, CM>(cf, p, d.sc, vv, vi);
        long long lp = 0;
        int slab = PML ? find_slab(ps, x, y, z, lp) : -1;
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            S* vp = V + n * d.sc + p;
            A v = T::load(vp, sc.v[n]);
            if (PML && slab >= 0) {
                const PmlSlab& b = ps.slab[slab];
                const long long q = lp + n * b.sc;
                A F = pp.vflux[q];
                A h = pp.vv[q] * v - pp.vvfo[q] * F;
                A t = F;
                t = t * vv[n];
                t = t + vi[n] * curl[n];
                pp.vflux[q] = t;
                v = h + pp.vvfn[q] * t;
            } else {
                v = v * vv[n];
                v = v + vi[n] * curl[n];
            }
            T::store(vp, v, sc.iv[n]);
        }
    }
}

template <typename S, int CM, bool PML>
__global__ void __launch_bounds__(256)
k_curr_march(S* __restrict__ I, const S* __restrict__ V, CoeffPtrs<typename StoreTraits<S>::A> cf, Dims d,
             Scales<typename StoreTraits<S>::A> sc, PmlSet ps, PmlPtrs<typename StoreTraits<S>::A> pp, MarchGrid g) {
    typedef StoreTraits<S> T;
    typedef typename T::A A;
    int tz, ty, cx;
    if (!march_tile(g, tz, ty, cx)) return;
    const int z = tz * 64 + threadIdx.x;
    const int y = ty * blockDim.y + threadIdx.y;
    if (z >= d.nz - 1 || y >= d.ny - 1) return;
    const int x0 = cx * g.chunk;
    const int x1 = min(d.nx - 1, x0 + g.chunk);
    if (x0 >= x1) return;
    const long long c1 = d.sc, c2 = 2 * d.sc;
    long long p = x0 * d.sx + y * d.sy + z;
    // V1, V2 of the current plane come from the previous iteration's x+1 loads
    A v1 = T::load(V + c1 + p, sc.v[1]);
    A v2 = T::load(V + c2 + p, sc.v[2]);
    for (int x = x0; x < x1; ++x, p += d.sx) {
        const A v0 = T::load(V + p, sc.v[0]);
        const A v2_xp = T::load(V + c2 + p + d.sx, sc.v[2]);
        const A v1_xp = T::load(V + c1 + p + d.sx, sc.v[1]);
        const A v2_yp = T::load(V + c2 + p + d.sy, sc.v[2]);
        const A v0_yp = T::load(V + p + d.sy, sc.v[0]);
        const A v1_zp = T::load(V + c1 + p + 1, sc.v[1]);
        const A v0_zp = T::load(V + p + 1, sc.v[0]);
        A curl[3];
        curl[0] = ((v2 - v2_yp) - v1) + v1_zp;
        curl[1] = ((v0 - v0_zp) - v2) + v2_xp;
        curl[2] = ((v1 - v1_xp) - v0) + v0_yp;
        v1 = v1_xp;
        v2 = v2_xp;

        A ii[3], iv[3];
        load_coeffs<A, CM>(cf, p, d.sc, ii, iv);
        long long lp = 0;
        int slab = PML ? find_slab(ps, x, y, z, lp) : -1;
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            S* ip = I + n * d.sc + p;
            A c = T::load(ip, sc.c[n]);
            if (PML && slab >= 0) {
                const PmlSlab& b = ps.slab[slab];
                const long long q = lp + n * b.sc;
                A F = pp.cflux[q];
                A h = pp.ii[q] * c - pp.iifo[q] * F;
                A t = F;
                t = t * ii[n];
                t = t + iv[n] * curl[n];
                pp.cflux[q] = t;
                c = h + pp.iifn[q] * t;
            } else {
                c = c * ii[n];
                c = c + iv[n] * curl[n];
            }
            T::store(ip, c, sc.ic[n]);
        }
    }
}

// ---------------------------------------------------------------------------
// sparse hooks
// ---------------------------------------------------------------------------
struct ExcDev {
    int count;
    int length;
    int period;                    // timesteps, 0 = none
    const long long* off;          // element offset incl. component
    const int* comp;
    const int* delay;
    const void* amp;               // A[count]
    const void* sig;               // A[length]
    int sequential;                // duplicate edges present: apply in list order
};

// Engine_Ext_Excitation::Apply2{Voltages,Current}Impl, engine_ext_excitation.cpp:33-94
template <typename S>
__global__ void k_excite(S* F, const unsigned* __restrict__ step, ExcDev e, Scales<typename StoreTraits<S>::A> sc, int is_curr) {
    typedef StoreTraits<S> T;
    typedef typename T::A A;
    const int numTS = (int)*step;
    const A* amp = (const A*)e.amp;
    const A* sig = (const A*)e.sig;
    int p = numTS + 1;
    if (e.period > 0) p = e.period;
    int begin = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    if (e.sequential) { if (begin != 0) return; stride = 1; }
    for (int i = begin; i < e.count; i += stride) {
        int k = numTS - e.delay[i];
        k *= (k > 0);
        k %= p;
        k *= (k < e.length);
        const int n = e.comp[i];
        const A s = is_curr ? sc.c[n] : sc.v[n];
        const A is = is_curr ? sc.ic[n] : sc.iv[n];
        S* fp = F + e.off[i];
        A v = T::load(fp, s);
        v = v + amp[i] * sig[k];
        T::store(fp, v, is);
    }
}

struct AdeDev {
    int count;
    const long long* off;          // [count] element offset of position (component 0)
    const void* v_int; const void* v_ext; const void* v_lor;   // [3][count]
    const void* i_int; const void* i_ext; const void* i_lor;
    void* volt_ade; void* volt_lor; void* curr_ade; void* curr_lor;  // state [3][count]
    int volt_on, curr_on, volt_lor_on, curr_lor_on;
};

// Engine_Ext_LorentzMaterial::DoPre{Voltage,Current}UpdatesImpl, engine_ext_lorentzmaterial.cpp:79-168
template <typename S>
__global__ void k_ade_pre(const S* __restrict__ F, AdeDev a, long long sc_field, Scales<typename StoreTraits<S>::A> sc, int is_curr) {
    typedef StoreTraits<S> T;
    typedef typename T::A A;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.count) return;
    const A* kint = (const A*)(is_curr ? a.i_int : a.v_int);
    const A* kext = (const A*)(is_curr ? a.i_ext : a.v_ext);
    const A* klor = (const A*)(is_curr ? a.i_lor : a.v_lor);
    A* J = (A*)(is_curr ? a.curr_ade : a.volt_ade);
    A* JL = (A*)(is_curr ? a.curr_lor : a.volt_lor);
    const int lor_on = is_curr ? a.curr_lor_on : a.volt_lor_on;
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        const long long q = (long long)n * a.count + i;
        const A f = T::load(F + n * sc_field + a.off[i], is_curr ? sc.c[n] : sc.v[n]);
        if (lor_on) {
            A jl = JL[q];
            A j = J[q];
            jl = jl + klor[q] * j;
            j = j * kint[q];
            j = j + kext[q] * (f - jl);
            JL[q] = jl;
            J[q] = j;
        } else {
            A j = J[q];
            j = j * kint[q];
            j = j + kext[q] * f;
            J[q] = j;
        }
    }
}

// Engine_Ext_Dispersive::Apply2{Voltages,Current}Impl, engine_ext_dispersive.cpp:76-127
template <typename S>
__global__ void k_ade_apply(S* F, AdeDev a, long long sc_field, Scales<typename StoreTraits<S>::A> sc, int is_curr) {
    typedef StoreTraits<S> T;
    typedef typename T::A A;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.count) return;
    const A* J = (const A*)(is_curr ? a.curr_ade : a.volt_ade);
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        S* fp = F + n * sc_field + a.off[i];
        const A s = is_curr ? sc.c[n] : sc.v[n];
        A f = T::load(fp, s);
        f = f - J[(long long)n * a.count + i];
        T::store(fp, f, is_curr ? sc.ic[n] : sc.iv[n]);
    }
}

struct TfsfDev {
    int count;
    int length, period, max_delay;
    const long long* off;   // element offset incl. component
    const int* comp;
    const int* delay;
    const void* delta;      // A[count]
    const void* amp;        // A[count]
    const void* sig;        // A[length] (curr signal for volt injection and vice versa)
};

// Engine_Ext_TFSF::DoPost{Voltage,Current}Updates, engine_ext_tfsf.cpp:36-215:
// F += (1-delta)*amp*s[L[d]] + delta*amp*s[L[d+1]] evaluated in double and
// rounded once to the storage type.
template <typename S>
__global__ void k_tfsf(S* F, const unsigned* __restrict__ step, TfsfDev t, Scales<typename StoreTraits<S>::A> sc, int is_curr) {
    typedef StoreTraits<S> T;
    typedef typename T::A A;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= t.count) return;
    const unsigned numTS = *step;
    const A* sig = (const A*)t.sig;
    const A* amp = (const A*)t.amp;
    const A* delta = (const A*)t.delta;
    auto lookup = [&](unsigned n) -> unsigned {
        unsigned L;
        if (numTS < n) L = 0;
        else if ((numTS - n >= (unsigned)t.length) && (t.period == 0)) L = 0;
        else L = numTS - n;
        if (t.period > 0) L = L % (unsigned)t.period;
        return L;
    };
    const unsigned dl = (unsigned)t.delay[i];
    const int n = t.comp[i];
    S* fp = F + t.off[i];
    const A s = is_curr ? sc.c[n] : sc.v[n];
    const double f = (double)T::load(fp, s);
    // (1.0-delta)*amp*s is double arithmetic; delta*amp*s is evaluated in the
    // storage float type, as in the reference expression
    const A second = delta[i] * amp[i] * sig[lookup(dl + 1)];
    double r = f + (1.0 - (double)delta[i]) * (double)amp[i] * (double)sig[lookup(dl)] + (double)second;
    T::store(fp, (A)r, is_curr ? sc.ic[n] : sc.iv[n]);
}

// ---------------------------------------------------------------------------
// step end: ++numTS, then probe integrals (FP64) and FD accumulation,
// ProcessIntegral::Process (Common/processintegral.cpp:134-171) at the new
// timestep, time = (numTS + 0.5*dual)*dT (FDTD/engine_interface_fdtd.h:52).
// ---------------------------------------------------------------------------
struct ProbeDev {
    int kind;              // 0 volt, 1 curr
    int first, count;      // edge range
    int td_interval, fd_interval;
    int n_freq;
    double weight;
    double t_shift;        // 0.5 for dual time
    double* td;            // [slots][2] (t, value)
    long long td_slots;
    double* fd;            // [n_freq][2]
    const double* freqs;
};

template <typename S>
__global__ void k_step_end(const S* __restrict__ V, const S* __restrict__ I, unsigned* step, int advance,
                           const ProbeDev* __restrict__ probes, int nprobes, const long long* __restrict__ e_off,
                           const int* __restrict__ e_comp, const signed char* __restrict__ e_sign,
                           Scales<typename StoreTraits<S>::A> sc, double dT) {
    typedef StoreTraits<S> T;
    __shared__ unsigned s_ts;
    if (threadIdx.x == 0) s_ts = *step + (unsigned)advance;
    __syncthreads();
    const unsigned ts = s_ts;
    for (int pi = threadIdx.x; pi < nprobes; pi += blockDim.x) {
        const ProbeDev& pr = probes[pi];
        const bool td = pr.td_interval > 0 && (ts % (unsigned)pr.td_interval) == 0;
        const bool fd = pr.fd_interval > 0 && (ts % (unsigned)pr.fd_interval) == 0;
        if (!td && !fd) continue;
        const S* F = pr.kind == 0 ? V : I;
        double val = 0.0;
        for (int e = pr.first; e < pr.first + pr.count; ++e) {
            const int n = e_comp[e];
            const double f = (double)T::load(F + e_off[e], pr.kind == 0 ? sc.v[n] : sc.c[n]);
            if (e_sign[e] > 0) val += f; else val -= f;
        }
        const double t = ((double)ts + pr.t_shift) * dT;
        const double w = val * pr.weight;
        if (td) {
            long long slot = ts / (unsigned)pr.td_interval;
            if (slot < pr.td_slots) { pr.td[2 * slot] = t; pr.td[2 * slot + 1] = w; }
        }
        if (fd) {
            for (int k = 0; k < pr.n_freq; ++k) {
                double sn, cs;
                sincos(-2.0 * M_PI * pr.freqs[k] * t, &sn, &cs);
                const double g = w * 2.0 * dT * (double)pr.fd_interval;
                pr.fd[2 * k] += g * cs;
                pr.fd[2 * k + 1] += g * sn;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) *step = ts;
}

// ---------------------------------------------------------------------------
// energy: eps0*sum(V^2) + mu0*sum(I^2) over x<Nx-1, y<Ny-1, all z
// (Engine_Interface_SSE_FDTD::CalcFastEnergy, engine_interface_sse_fdtd.cpp:37-69),
// accumulated in FP64, deterministic two-pass reduction.
// ---------------------------------------------------------------------------
template <typename S>
__global__ void __launch_bounds__(256)
k_energy_partial(const S* __restrict__ V, const S* __restrict__ I, Dims d, Scales<typename StoreTraits<S>::A> sc, double* part) {
    typedef StoreTraits<S> T;
    __shared__ double se[256], sh[256];
    double e = 0.0, h = 0.0;
    const long long rows = (long long)(d.nx - 1) * (d.ny - 1);
    for (long long r = blockIdx.x; r < rows; r += gridDim.x) {
        const int x = (int)(r / (d.ny - 1)), y = (int)(r % (d.ny - 1));
        const long long base = x * d.sx + y * d.sy;
        for (int z = threadIdx.x; z < d.nz; z += blockDim.x) {
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                const double v = (double)T::load(V + n * d.sc + base + z, sc.v[n]);
                const double c = (double)T::load(I + n * d.sc + base + z, sc.c[n]);
                e += v * v;
                h += c * c;
            }
        }
    }
    se[threadIdx.x] = e;
    sh[threadIdx.x] = h;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) { se[threadIdx.x] += se[threadIdx.x + s]; sh[threadIdx.x] += sh[threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = se[0]; part[2 * blockIdx.x + 1] = sh[0]; }
}

__global__ void k_energy_final(const double* part, int n, double* out) {
    __shared__ double se[256], sh[256];
    double e = 0.0, h = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) { e += part[2 * i]; h += part[2 * i + 1]; }
    se[threadIdx.x] = e;
    sh[threadIdx.x] = h;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) { se[threadIdx.x] += se[threadIdx.x + s]; sh[threadIdx.x] += sh[threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = se[0]; out[1] = sh[0]; }
}

// plain streaming copy: the calibration kernel for HBM byte counters and the
// achievable-bandwidth reference
__global__ void __launch_bounds__(256) k_copy(const float4* __restrict__ a, float4* __restrict__ b, long long n) {
    // four independent 16-byte loads in flight per lane, then four stores
    const long long stride = (long long)gridDim.x * blockDim.x;
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const float4 r0 = a[i], r1 = a[i + stride], r2 = a[i + 2 * stride], r3 = a[i + 3 * stride];
        b[i] = r0; b[i + stride] = r1; b[i + 2 * stride] = r2; b[i + 3 * stride] = r3;
    }
    for (; i < n; i += stride) b[i] = a[i];
}

// gather/scatter between the padded device layout and the dense host layout
template <typename S, typename H>
__global__ void k_pack(const S* __restrict__ F, H* __restrict__ out, Dims d, Scales<typename StoreTraits<S>::A> sc, int is_curr) {
    typedef StoreTraits<S> T;
    const long long total = 3LL * d.nx * d.ny * d.nz;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
        long long r = i;
        const int z = (int)(r % d.nz); r /= d.nz;
        const int y = (int)(r % d.ny); r /= d.ny;
        const int x = (int)(r % d.nx); const int n = (int)(r / d.nx);
        out[i] = (H)T::load(F + n * d.sc + x * d.sx + y * d.sy + z, is_curr ? sc.c[n] : sc.v[n]);
    }
}

template <typename S, typename H>
__global__ void k_unpack(S* __restrict__ F, const H* __restrict__ in, Dims d, Scales<typename StoreTraits<S>::A> sc, int is_curr) {
    typedef StoreTraits<S> T;
    typedef typename T::A A;
    const long long total = 3LL * d.nx * d.ny * d.nz;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
        long long r = i;
        const int z = (int)(r % d.nz); r /= d.nz;
        const int y = (int)(r % d.ny); r /= d.ny;
        const int x = (int)(r % d.nx); const int n = (int)(r / d.nx);
        T::store(F + n * d.sc + x * d.sx + y * d.sy + z, (A)in[i], is_curr ? sc.ic[n] : sc.iv[n]);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    int alloc(size_t n) {
        if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
        if (n == 0) return FDTD_OK;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) { p = nullptr; return fail(FDTD_ERR_NOMEM, "hipMalloc(%zu) failed: %s", n, hipGetErrorString(e)); }
        bytes = n;
        return FDTD_OK;
    }
    template <typename X> X* as() const { return (X*)p; }
};

struct ProbeHost {
    int kind;
    int norm_dir;
    double weight;
    uint32_t td_interval, fd_interval;
    std::vector<double> freqs;
    std::vector<long long> off;
    std::vector<int> comp;
    std::vector<signed char> sign;
    DevBuf td, fd, dfreq;
    long long td_slots = 0;
};

struct AdeHost {
    uint32_t count = 0;
    uint32_t flags = 0;
    DevBuf off, coef, state;   // coef: 6 lists [3][count]; state: 4 lists [3][count]
};

} // namespace

struct fdtd_engine {
    int prec = FDTD_PREC_FP32;
    int dev = 0;
    uint32_t N[3] = {0, 0, 0};
    Dims d{};
    double dT = 0;
    size_t elem = 4;              // bytes per stored field value
    size_t acc = 4;               // bytes per arithmetic value (coefficients)
    hipStream_t stream = nullptr;

    DevBuf V, I;                  // fields
    DevBuf step;                  // unsigned numTS on device
    uint32_t numTS = 0;

    int cf_mode = -1;             // CF_* once set
    DevBuf cf_vv, cf_vi, cf_ii, cf_iv;   // full mode
    DevBuf cf_idx, cf_tabE, cf_tabH;     // compressed mode
    uint32_t n_entries = 0;

    double vscale[3] = {1, 1, 1}, cscale[3] = {1, 1, 1};

    // UPML
    PmlSet pml{};
    DevBuf pml_coef;              // 6 arrays over all slabs
    DevBuf pml_flux;              // 2 arrays over all slabs
    std::vector<std::vector<char>> pml_host_coef;   // staged per slab until finalized
    long long pml_total = 0;
    bool pml_dirty = false;
    bool pml_final = false;

    // excitation
    bool has_exc = false;
    ExcDev exc_v{}, exc_c{};
    DevBuf exc_buf;

    // ADE orders
    std::vector<AdeHost*> ade;

    // TFSF
    TfsfDev tfsf_v{}, tfsf_c{};
    DevBuf tfsf_buf_v, tfsf_buf_c;

    // probes
    std::vector<ProbeHost*> probes;
    DevBuf probe_dev, e_off, e_comp, e_sign;
    bool probes_dirty = true;
    bool sampled0 = false;

    // energy
    DevBuf e_part, e_out;
    int e_blocks = 0;

    // graph
    bool use_graph = true;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_steps = 0;
    hipGraphExec_t gexec1 = nullptr;
    hipGraph_t graph1 = nullptr;

    // kernel selection: 0 = one thread per cell, 1 = x-march, 2 = banded, 3 = vectors (default)
    int variant = 3;
    int tile_y = 4;
    int chunk_x = 64;

    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev;
    size_t ev_used = 0;
    double t_volt = 0, t_curr = 0;
    uint32_t n_volt = 0, n_curr = 0;

    ~fdtd_engine() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (gexec1) (void)hipGraphExecDestroy(gexec1);
        if (graph1) (void)hipGraphDestroy(graph1);
        for (auto e : ev) (void)hipEventDestroy(e);
        for (auto p : probes) delete p;
        for (auto a : ade) delete a;
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void invalidate_graph(fdtd_engine* h) {
    if (h->gexec) { (void)hipGraphExecDestroy(h->gexec); h->gexec = nullptr; }
    if (h->graph) { (void)hipGraphDestroy(h->graph); h->graph = nullptr; }
    if (h->gexec1) { (void)hipGraphExecDestroy(h->gexec1); h->gexec1 = nullptr; }
    if (h->graph1) { (void)hipGraphDestroy(h->graph1); h->graph1 = nullptr; }
    h->graph_steps = 0;
}

template <typename A>
Scales<A> make_scales(const fdtd_engine* h) {
    Scales<A> s;
    for (int n = 0; n < 3; ++n) {
        s.v[n] = (A)h->vscale[n]; s.iv[n] = (A)(1.0 / h->vscale[n]);
        s.c[n] = (A)h->cscale[n]; s.ic[n] = (A)(1.0 / h->cscale[n]);
    }
    return s;
}

long long cell_off(const fdtd_engine* h, uint32_t x, uint32_t y, uint32_t z) {
    return (long long)x * h->d.sx + (long long)y * h->d.sy + z;
}

// convert a host array of dtype into the handle's arithmetic type
int convert_host(const fdtd_engine* h, int dtype, const void* src, size_t n, std::vector<char>& dst) {
    if (!src) return fail(FDTD_ERR_ARG, "null array");
    if (dtype != FDTD_DTYPE_F32 && dtype != FDTD_DTYPE_F64) return fail(FDTD_ERR_ARG, "bad dtype %d", dtype);
    dst.resize(n * h->acc);
    if (h->acc == 8) {
        double* o = (double*)dst.data();
        if (dtype == FDTD_DTYPE_F64) memcpy(o, src, n * 8);
        else for (size_t i = 0; i < n; ++i) o[i] = (double)((const float*)src)[i];
    } else {
        float* o = (float*)dst.data();
        if (dtype == FDTD_DTYPE_F32) memcpy(o, src, n * 4);
        else for (size_t i = 0; i < n; ++i) o[i] = (float)((const double*)src)[i];
    }
    return FDTD_OK;
}

// dense [3][Nx][Ny][Nz] host array -> padded device array in arithmetic type
int upload_dense3(fdtd_engine* h, int dtype, const void* src, DevBuf& buf) {
    const size_t nd = (size_t)h->N[0] * h->N[1] * h->N[2];
    int rc = buf.alloc((size_t)3 * h->d.sc * h->acc);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(buf.p, 0, buf.bytes, h->stream));
    std::vector<char> row;
    const size_t in_elem = dtype == FDTD_DTYPE_F64 ? 8 : 4;
    // copy plane by plane with the strided 2D copy
    std::vector<char> tmp;
    for (int n = 0; n < 3; ++n) {
        const char* s = (const char*)src + (size_t)n * nd * in_elem;
        rc = convert_host(h, dtype, s, nd, tmp);
        if (rc) return rc;
        HIPCHK(hipMemcpy2DAsync((char*)buf.p + (size_t)n * h->d.sc * h->acc, (size_t)h->d.sy * h->acc,
                                tmp.data(), (size_t)h->N[2] * h->acc, (size_t)h->N[2] * h->acc,
                                (size_t)h->N[0] * h->N[1], hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    return FDTD_OK;
}

template <typename S>
int launch_step_kernels(fdtd_engine* h, bool timed);

int capture_graph(fdtd_engine* h, int steps, hipGraph_t* g, hipGraphExec_t* ge);

int finalize_pml(fdtd_engine* h) {
    if (!h->pml_dirty) return FDTD_OK;
    const size_t a = h->acc;
    int rc = h->pml_coef.alloc((size_t)6 * h->pml_total * a);
    if (rc) return rc;
    rc = h->pml_flux.alloc((size_t)2 * h->pml_total * a);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(h->pml_flux.p, 0, h->pml_flux.bytes, h->stream));
    for (int s = 0; s < h->pml.count; ++s) {
        const PmlSlab& b = h->pml.slab[s];
        const size_t cnt = (size_t)3 * b.sc;
        for (int k = 0; k < 6; ++k) {
            const char* src = h->pml_host_coef[s].data() + (size_t)k * cnt * a;
            HIPCHK(hipMemcpyAsync((char*)h->pml_coef.p + ((size_t)k * h->pml_total + b.off) * a, src, cnt * a,
                                  hipMemcpyHostToDevice, h->stream));
        }
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    h->pml_host_coef.clear();
    h->pml_dirty = false;
    h->pml_final = true;
    return FDTD_OK;
}

int finalize_probes(fdtd_engine* h, uint32_t upto_ts) {
    // (re)build the device probe table; td buffers sized for samples up to upto_ts
    std::vector<ProbeDev> pd;
    std::vector<long long> off;
    std::vector<int> comp;
    std::vector<signed char> sign;
    bool realloc_needed = h->probes_dirty;
    for (auto p : h->probes) {
        if (p->td_interval > 0 && (long long)(upto_ts / p->td_interval) + 1 > p->td_slots) realloc_needed = true;
    }
    if (!realloc_needed) return FDTD_OK;
    for (auto p : h->probes) {
        ProbeDev q{};
        q.kind = p->kind;
        q.first = (int)off.size();
        q.count = (int)p->off.size();
        q.td_interval = (int)p->td_interval;
        q.fd_interval = (int)p->fd_interval;
        q.n_freq = (int)p->freqs.size();
        q.weight = p->weight;
        q.t_shift = p->kind == FDTD_PROBE_CURRENT ? 0.5 : 0.0;
        if (p->td_interval > 0) {
            long long need = (long long)(upto_ts / p->td_interval) + 1;
            if (need > p->td_slots) {
                long long ns = std::max(need, 2 * p->td_slots);
                DevBuf nb;
                int rc = nb.alloc((size_t)ns * 2 * sizeof(double));
                if (rc) return rc;
                HIPCHK(hipMemsetAsync(nb.p, 0xff, nb.bytes, h->stream));   // NaN fill = not sampled
                if (p->td.p) HIPCHK(hipMemcpyAsync(nb.p, p->td.p, (size_t)p->td_slots * 2 * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
                HIPCHK(hipStreamSynchronize(h->stream));
                std::swap(p->td.p, nb.p);
                std::swap(p->td.bytes, nb.bytes);
                p->td_slots = ns;
            }
        }
        q.td = p->td.as<double>();
        q.td_slots = p->td_slots;
        q.fd = p->fd.as<double>();
        q.freqs = p->dfreq.as<double>();
        pd.push_back(q);
        off.insert(off.end(), p->off.begin(), p->off.end());
        comp.insert(comp.end(), p->comp.begin(), p->comp.end());
        sign.insert(sign.end(), p->sign.begin(), p->sign.end());
    }
    int rc;
    if ((rc = h->probe_dev.alloc(std::max<size_t>(1, pd.size()) * sizeof(ProbeDev)))) return rc;
    if ((rc = h->e_off.alloc(std::max<size_t>(1, off.size()) * sizeof(long long)))) return rc;
    if ((rc = h->e_comp.alloc(std::max<size_t>(1, comp.size()) * sizeof(int)))) return rc;
    if ((rc = h->e_sign.alloc(std::max<size_t>(1, sign.size())))) return rc;
    if (!pd.empty()) HIPCHK(hipMemcpyAsync(h->probe_dev.p, pd.data(), pd.size() * sizeof(ProbeDev), hipMemcpyHostToDevice, h->stream));
    if (!off.empty()) {
        HIPCHK(hipMemcpyAsync(h->e_off.p, off.data(), off.size() * sizeof(long long), hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(h->e_comp.p, comp.data(), comp.size() * sizeof(int), hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(h->e_sign.p, sign.data(), sign.size(), hipMemcpyHostToDevice, h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    h->probes_dirty = false;
    invalidate_graph(h);   // kernel arguments changed
    return FDTD_OK;
}

template <typename S>
int launch_step_end(fdtd_engine* h, int advance) {
    typedef typename StoreTraits<S>::A A;
    hipLaunchKernelGGL((k_step_end<S>), dim3(1), dim3(256), 0, h->stream, h->V.as<S>(), h->I.as<S>(), h->step.as<unsigned>(), advance,
                       h->probe_dev.as<ProbeDev>(), (int)h->probes.size(), h->e_off.as<long long>(), h->e_comp.as<int>(),
                       h->e_sign.as<signed char>(), make_scales<A>(h), h->dT);
    return FDTD_OK;
}

template <typename S, bool PML>
void launch_volt_cm(fdtd_engine* h, dim3 g, dim3 b, CoeffPtrs<typename StoreTraits<S>::A> cf, PmlPtrs<typename StoreTraits<S>::A> pp) {
    typedef typename StoreTraits<S>::A A;
    Scales<A> sc = make_scales<A>(h);
    switch (h->cf_mode) {
    case CF_FULL: hipLaunchKernelGGL((k_volt<S, CF_FULL, PML>), g, b, 0, h->stream, h->V.as<S>(), h->I.as<S>(), cf, h->d, sc, h->pml, pp); break;
    case CF_IDX16: hipLaunchKernelGGL((k_volt<S, CF_IDX16, PML>), g, b, 0, h->stream, h->V.as<S>(), h->I.as<S>(), cf, h->d, sc, h->pml, pp); break;
    default: hipLaunchKernelGGL((k_volt<S, CF_IDX32, PML>), g, b, 0, h->stream, h->V.as<S>(), h->I.as<S>(), cf, h->d, sc, h->pml, pp); break;
    }
}

template <typename S, bool PML>
void launch_curr_cm(fdtd_engine* h, dim3 g, dim3 b, CoeffPtrs<typename StoreTraits<S>::A> cf, PmlPtrs<typename StoreTraits<S>::A> pp) {
    typedef typename StoreTraits<S>::A A;
    Scales<A> sc = make_scales<A>(h);
    switch (h->cf_mode) {
    case CF_FULL: hipLaunchKernelGGL((k_curr<S, CF_FULL, PML>), g, b, 0, h->stream, h->I.as<S>(), h->V.as<S>(), cf, h->d, sc, h->pml, pp); break;
    case CF_IDX16: hipLaunchKernelGGL((k_curr<S, CF_IDX16, PML>), g, b, 0, h->stream, h->I.as<S>(), h->V.as<S>(), cf, h->d, sc, h->pml, pp); break;
    default: hipLaunchKernelGGL((k_curr<S, CF_IDX32, PML>), g, b, 0, h->stream, h->I.as<S>(), h->V.as<S>(), cf, h->d, sc, h->pml, pp); break;
    }
}

MarchGrid make_march(const fdtd_engine* h, int ny_eff, int nx_eff) {
    MarchGrid g;
    g.tiles_z = (h->N[2] + 63) / 64;
    g.tiles_y = (ny_eff + h->tile_y - 1) / h->tile_y;
    g.chunk = std::max(1, std::min(h->chunk_x, nx_eff));
    g.chunks_x = (nx_eff + g.chunk - 1) / g.chunk;
    g.nblocks = g.tiles_z * g.tiles_y * g.chunks_x;
    g.per_xcd = (g.nblocks + 7) / 8;
    return g;
}

template <typename S, bool PML>
void launch_volt_march(fdtd_engine* h, CoeffPtrs<typename StoreTraits<S>::A> cf, PmlPtrs<typename StoreTraits<S>::A> pp) {
    typedef typename StoreTraits<S>::A A;
    Scales<A> sc = make_scales<A>(h);
    MarchGrid g = make_march(h, h->N[1], h->N[0]);
    const dim3 b(64, h->tile_y, 1), gr(g.per_xcd * 8);
    switch (h->cf_mode) {
    case CF_FULL: hipLaunchKernelGGL((k_volt_march<S, CF_FULL, PML>), gr, b, 0, h->stream, h->V.as<S>(), h->I.as<S>(), cf, h->d, sc, h->pml, pp, g); break;
    case CF_IDX16: hipLaunchKernelGGL((k_volt_march<S, CF_IDX16, PML>), gr, b, 0, h->stream, h->V.as<S>(), h->I.as<S>(), cf, h->d, sc, h->pml, pp, g); break;
    default: hipLaunchKernelGGL((k_volt_march<S, CF_IDX32, PML>), gr, b, 0, h->stream, h->V.as<S>(), h->I.as<S>(), cf, h->d, sc, h->pml, pp, g); break;
    }
}

template <typename S, bool PML>
void launch_curr_march(fdtd_engine* h, CoeffPtrs<typename StoreTraits<S>::A> cf, PmlPtrs<typename StoreTraits<S>::A> pp) {
    typedef typename StoreTraits<S>::A A;
    Scales<A> sc = make_scales<A>(h);
    MarchGrid g = make_march(h, h->N[1] - 1, h->N[0] - 1);
    const dim3 b(64, h->tile_y, 1), gr(g.per_xcd * 8);
    switch (h->cf_mode) {
    case CF_FULL: hipLaunchKernelGGL((k_curr_march<S, CF_FULL, PML>), gr, b, 0, h->stream, h->I.as<S>(), h->V.as<S>(), cf, h->d, sc, h->pml, pp, g); break;
    case CF_IDX16: hipLaunchKernelGGL((k_curr_march<S, CF_IDX16, PML>), gr, b, 0, h->stream, h->I.as<S>(), h->V.as<S>(), cf, h->d, sc, h->pml, pp, g); break;
    default: hipLaunchKernelGGL((k_curr_march<S, CF_IDX32, PML>), gr, b, 0, h->stream, h->I.as<S>(), h->V.as<S>(), cf, h->d, sc, h->pml, pp, g); break;
    }
}

template <typename S> struct VecWidth { enum { value = 16 / sizeof(S) }; };

template <typename S, bool PML>
void launch_volt_vec(fdtd_engine* h, CoeffPtrs<typename StoreTraits<S>::A> cf, PmlPtrs<typename StoreTraits<S>::A> pp) {
    typedef typename StoreTraits<S>::A A;
    constexpr int VZ = VecWidth<S>::value;
    Scales<A> sc = make_scales<A>(h);
    const long long plane = (long long)h->N[1] * h->d.sy;
    const dim3 b(256, 1, 1);
    const dim3 g((unsigned)((plane / VZ + 255) / 256), h->N[0], 1);
    switch (h->cf_mode) {
    case CF_FULL: hipLaunchKernelGGL((k_volt_vec<S, CF_FULL, PML, VZ>), g, b, 0, h->stream, h->V.as<S>(), h->I.as<S>(), cf, h->d, sc, h->pml, pp); break;
    case CF_IDX16: hipLaunchKernelGGL((k_volt_vec<S, CF_IDX16, PML, VZ>), g, b, 0, h->stream, h->V.as<S>(), h->I.as<S>(), cf, h->d, sc, h->pml, pp); break;
    default: hipLaunchKernelGGL((k_volt_vec<S, CF_IDX32, PML, VZ>), g, b, 0, h->stream, h->V.as<S>(), h->I.as<S>(), cf, h->d, sc, h->pml, pp); break;
    }
}

template <typename S, bool PML>
void launch_curr_vec(fdtd_engine* h, CoeffPtrs<typename StoreTraits<S>::A> cf, PmlPtrs<typename StoreTraits<S>::A> pp) {
    typedef typename StoreTraits<S>::A A;
    constexpr int VZ = VecWidth<S>::value;
    Scales<A> sc = make_scales<A>(h);
    const long long plane = (long long)(h->N[1] - 1) * h->d.sy;
    const dim3 b(256, 1, 1);
    const dim3 g((unsigned)((plane / VZ + 255) / 256), h->N[0] > 1 ? h->N[0] - 1 : 1, 1);
    switch (h->cf_mode) {
    case CF_FULL: hipLaunchKernelGGL((k_curr_vec<S, CF_FULL, PML, VZ>), g, b, 0, h->stream, h->I.as<S>(), h->V.as<S>(), cf, h->d, sc, h->pml, pp); break;
    case CF_IDX16: hipLaunchKernelGGL((k_curr_vec<S, CF_IDX16, PML, VZ>), g, b, 0, h->stream, h->I.as<S>(), h->V.as<S>(), cf, h->d, sc, h->pml, pp); break;
    default: hipLaunchKernelGGL((k_curr_vec<S, CF_IDX32, PML, VZ>), g, b, 0, h->stream, h->I.as<S>(), h->V.as<S>(), cf, h->d, sc, h->pml, pp); break;
    }
}

BandGrid make_band(const fdtd_engine* h, int ny_eff, int nx_eff) {
    BandGrid g;
    g.tiles_z = (h->N[2] + 63) / 64;
    g.tiles_y = (ny_eff + h->tile_y - 1) / h->tile_y;
    g.band_tiles = std::max(1, std::min(h->chunk_x, g.tiles_y));
    g.nbands = (g.tiles_y + g.band_tiles - 1) / g.band_tiles;
    g.nx = nx_eff;
    g.nblocks = g.tiles_z * g.band_tiles * g.nx * g.nbands;
    g.per_xcd = (g.nblocks + 7) / 8;
    return g;
}

template <typename S, bool PML>
void launch_volt_band(fdtd_engine* h, CoeffPtrs<typename StoreTraits<S>::A> cf, PmlPtrs<typename StoreTraits<S>::A> pp) {
    typedef typename StoreTraits<S>::A A;
    Scales<A> sc = make_scales<A>(h);
    BandGrid g = make_band(h, h->N[1], h->N[0]);
    const dim3 b(64, h->tile_y, 1), gr(g.per_xcd * 8);
    switch (h->cf_mode) {
    case CF_FULL: hipLaunchKernelGGL((k_volt_band<S, CF_FULL, PML>), gr, b, 0, h->stream, h->V.as<S>(), h->I.as<S>(), cf, h->d, sc, h->pml, pp, g); break;
    case CF_IDX16: hipLaunchKernelGGL((k_volt_band<S, CF_IDX16, PML>), gr, b, 0, h->stream, h->V.as<S>(), h->I.as<S>(), cf, h->d, sc, h->pml, pp, g); break;
    default: hipLaunchKernelGGL((k_volt_band<S, CF_IDX32, PML>), gr, b, 0, h->stream, h->V.as<S>(), h->I.as<S>(), cf, h->d, sc, h->pml, pp, g); break;
    }
}

template <typename S, bool PML>
void launch_curr_band(fdtd_engine* h, CoeffPtrs<typename StoreTraits<S>::A> cf, PmlPtrs<typename StoreTraits<S>::A> pp) {
    typedef typename StoreTraits<S>::A A;
    Scales<A> sc = make_scales<A>(h);
    BandGrid g = make_band(h, h->N[1] - 1, h->N[0] - 1);
    const dim3 b(64, h->tile_y, 1), gr(g.per_xcd * 8);
    switch (h->cf_mode) {
    case CF_FULL: hipLaunchKernelGGL((k_curr_band<S, CF_FULL, PML>), gr, b, 0, h->stream, h->I.as<S>(), h->V.as<S>(), cf, h->d, sc, h->pml, pp, g); break;
    case CF_IDX16: hipLaunchKernelGGL((k_curr_band<S, CF_IDX16, PML>), gr, b, 0, h->stream, h->I.as<S>(), h->V.as<S>(), cf, h->d, sc, h->pml, pp, g); break;
    default: hipLaunchKernelGGL((k_curr_band<S, CF_IDX32, PML>), gr, b, 0, h->stream, h->I.as<S>(), h->V.as<S>(), cf, h->d, sc, h->pml, pp, g); break;
    }
}

template <typename S>
int launch_step_kernels(fdtd_engine* h, bool timed) {
    typedef typename StoreTraits<S>::A A;
    Scales<A> sc = make_scales<A>(h);
    const dim3 blk(64, 4, 1);
    const dim3 grd((h->N[2] + 63) / 64, (h->N[1] + 3) / 4, h->N[0]);
    const dim3 grdc((h->N[2] + 63) / 64, (h->N[1] + 3) / 4, h->N[0] > 1 ? h->N[0] - 1 : 1);

    CoeffPtrs<A> cfe{}, cfh{};
    if (h->cf_mode == CF_FULL) {
        cfe.a = h->cf_vv.as<A>(); cfe.b = h->cf_vi.as<A>();
        cfh.a = h->cf_ii.as<A>(); cfh.b = h->cf_iv.as<A>();
    } else {
        cfe.a = h->cf_tabE.as<A>(); cfe.idx = h->cf_idx.p;
        cfh.a = h->cf_tabH.as<A>(); cfh.idx = h->cf_idx.p;
    }
    PmlPtrs<A> pp{};
    if (h->pml.count) {
        A* c = h->pml_coef.as<A>();
        const long long t = h->pml_total;
        pp.vv = c; pp.vvfo = c + t; pp.vvfn = c + 2 * t;
        pp.ii = c + 3 * t; pp.iifo = c + 4 * t; pp.iifn = c + 5 * t;
        pp.vflux = h->pml_flux.as<A>(); pp.cflux = h->pml_flux.as<A>() + t;
    }
    const bool pml = h->pml.count > 0;
    const unsigned* stp = h->step.as<unsigned>();

    // ---- voltage half step ----
    auto ade_dev = [&](AdeHost* a) {
        AdeDev ad{};
        const size_t n3 = (size_t)3 * a->count;
        A* cf = a->coef.as<A>();
        A* st = a->state.as<A>();
        ad.count = (int)a->count;
        ad.off = a->off.as<long long>();
        ad.v_int = cf; ad.v_ext = cf + n3; ad.v_lor = cf + 2 * n3;
        ad.i_int = cf + 3 * n3; ad.i_ext = cf + 4 * n3; ad.i_lor = cf + 5 * n3;
        ad.volt_ade = st; ad.volt_lor = st + n3; ad.curr_ade = st + 2 * n3; ad.curr_lor = st + 3 * n3;
        ad.volt_on = (a->flags & 1) != 0; ad.curr_on = (a->flags & 2) != 0;
        ad.volt_lor_on = (a->flags & 4) != 0; ad.curr_lor_on = (a->flags & 8) != 0;
        return ad;
    };
    for (auto a : h->ade)
        if ((a->flags & 1) && a->count)
            hipLaunchKernelGGL((k_ade_pre<S>), dim3((a->count + 255) / 256), dim3(256), 0, h->stream, h->V.as<S>(), ade_dev(a), h->d.sc, sc, 0);
    const size_t evi = timed ? h->ev_used : 0;
    if (timed) h->ev_used += 4;
    if (timed) HIPCHK(hipEventRecord(h->ev[evi], h->stream));
    if (h->variant == 3) { if (pml) launch_volt_vec<S, true>(h, cfe, pp); else launch_volt_vec<S, false>(h, cfe, pp); }
    else if (h->variant == 2) { if (pml) launch_volt_band<S, true>(h, cfe, pp); else launch_volt_band<S, false>(h, cfe, pp); }
    else if (h->variant == 1) { if (pml) launch_volt_march<S, true>(h, cfe, pp); else launch_volt_march<S, false>(h, cfe, pp); }
    else if (pml) launch_volt_cm<S, true>(h, grd, blk, cfe, pp); else launch_volt_cm<S, false>(h, grd, blk, cfe, pp);
    if (timed) HIPCHK(hipEventRecord(h->ev[evi + 1], h->stream));
    if (h->tfsf_v.count)
        hipLaunchKernelGGL((k_tfsf<S>), dim3((h->tfsf_v.count + 255) / 256), dim3(256), 0, h->stream, h->V.as<S>(), stp, h->tfsf_v, sc, 0);
    for (auto a : h->ade)
        if ((a->flags & 1) && a->count)
            hipLaunchKernelGGL((k_ade_apply<S>), dim3((a->count + 255) / 256), dim3(256), 0, h->stream, h->V.as<S>(), ade_dev(a), h->d.sc, sc, 0);
    if (h->has_exc && h->exc_v.count)
        hipLaunchKernelGGL((k_excite<S>), dim3(std::min(1024, (h->exc_v.count + 255) / 256)), dim3(256), 0, h->stream, h->V.as<S>(), stp, h->exc_v, sc, 0);

    // ---- current half step ----
    for (auto a : h->ade)
        if ((a->flags & 2) && a->count)
            hipLaunchKernelGGL((k_ade_pre<S>), dim3((a->count + 255) / 256), dim3(256), 0, h->stream, h->I.as<S>(), ade_dev(a), h->d.sc, sc, 1);
    if (timed) HIPCHK(hipEventRecord(h->ev[evi + 2], h->stream));
    if (h->variant == 3) { if (pml) launch_curr_vec<S, true>(h, cfh, pp); else launch_curr_vec<S, false>(h, cfh, pp); }
    else if (h->variant == 2) { if (pml) launch_curr_band<S, true>(h, cfh, pp); else launch_curr_band<S, false>(h, cfh, pp); }
    else if (h->variant == 1) { if (pml) launch_curr_march<S, true>(h, cfh, pp); else launch_curr_march<S, false>(h, cfh, pp); }
    else if (pml) launch_curr_cm<S, true>(h, grdc, blk, cfh, pp); else launch_curr_cm<S, false>(h, grdc, blk, cfh, pp);
    if (timed) HIPCHK(hipEventRecord(h->ev[evi + 3], h->stream));
    if (h->tfsf_c.count)
        hipLaunchKernelGGL((k_tfsf<S>), dim3((h->tfsf_c.count + 255) / 256), dim3(256), 0, h->stream, h->I.as<S>(), stp, h->tfsf_c, sc, 1);
    for (auto a : h->ade)
        if ((a->flags & 2) && a->count)
            hipLaunchKernelGGL((k_ade_apply<S>), dim3((a->count + 255) / 256), dim3(256), 0, h->stream, h->I.as<S>(), ade_dev(a), h->d.sc, sc, 1);
    if (h->has_exc && h->exc_c.count)
        hipLaunchKernelGGL((k_excite<S>), dim3(std::min(1024, (h->exc_c.count + 255) / 256)), dim3(256), 0, h->stream, h->I.as<S>(), stp, h->exc_c, sc, 1);

    return launch_step_end<S>(h, 1);
}

int launch_step(fdtd_engine* h, bool timed) {
    switch (h->prec) {
    case FDTD_PREC_FP32: return launch_step_kernels<float>(h, timed);
    case FDTD_PREC_FP64: return launch_step_kernels<double>(h, timed);
    case FDTD_PREC_FP16: return launch_step_kernels<__half>(h, timed);
    default: return launch_step_kernels<__hip_bfloat16>(h, timed);
    }
}

int capture_graph(fdtd_engine* h, int steps, hipGraph_t* g, hipGraphExec_t* ge) {
    HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < steps; ++s) {
        int rc = launch_step(h, false);
        if (rc) { hipGraph_t dummy; (void)hipStreamEndCapture(h->stream, &dummy); return rc; }
    }
    HIPCHK(hipStreamEndCapture(h->stream, g));
    HIPCHK(hipGraphInstantiate(ge, *g, nullptr, nullptr, 0));
    return FDTD_OK;
}

int ready_check(fdtd_engine* h) {
    if (h->cf_mode < 0) return fail(FDTD_ERR_STATE, "operator coefficients not set");
    return FDTD_OK;
}

template <typename S>
int energy_impl(fdtd_engine* h, double* out) {
    typedef typename StoreTraits<S>::A A;
    int rc;
    if (!h->e_part.p) {
        h->e_blocks = 2048;
        if ((rc = h->e_part.alloc((size_t)h->e_blocks * 2 * sizeof(double)))) return rc;
        if ((rc = h->e_out.alloc(2 * sizeof(double)))) return rc;
    }
    hipLaunchKernelGGL((k_energy_partial<S>), dim3(h->e_blocks), dim3(256), 0, h->stream, h->V.as<S>(), h->I.as<S>(), h->d, make_scales<A>(h), h->e_part.as<double>());
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, h->stream, h->e_part.as<double>(), h->e_blocks, h->e_out.as<double>());
    double r[2];
    HIPCHK(hipMemcpyAsync(r, h->e_out.p, sizeof(r), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    *out = kEPS0 * r[0] + kMUE0 * r[1];
    return FDTD_OK;
}

template <typename S, typename H>
int get_field_impl(fdtd_engine* h, int which, H* out) {
    typedef typename StoreTraits<S>::A A;
    const size_t n = (size_t)3 * h->N[0] * h->N[1] * h->N[2];
    DevBuf tmp;
    int rc = tmp.alloc(n * sizeof(H));
    if (rc) return rc;
    hipLaunchKernelGGL((k_pack<S, H>), dim3(2048), dim3(256), 0, h->stream, (which == FDTD_VOLT ? h->V : h->I).template as<S>(), tmp.as<H>(), h->d, make_scales<A>(h), which);
    HIPCHK(hipMemcpyAsync(out, tmp.p, n * sizeof(H), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return FDTD_OK;
}

template <typename S, typename H>
int set_field_impl(fdtd_engine* h, int which, const H* in) {
    typedef typename StoreTraits<S>::A A;
    const size_t n = (size_t)3 * h->N[0] * h->N[1] * h->N[2];
    DevBuf tmp;
    int rc = tmp.alloc(n * sizeof(H));
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(tmp.p, in, n * sizeof(H), hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL((k_unpack<S, H>), dim3(2048), dim3(256), 0, h->stream, (which == FDTD_VOLT ? h->V : h->I).template as<S>(), tmp.as<H>(), h->d, make_scales<A>(h), which);
    HIPCHK(hipStreamSynchronize(h->stream));
    return FDTD_OK;
}

template <typename H>
int field_dispatch_get(fdtd_engine* h, int which, H* out) {
    switch (h->prec) {
    case FDTD_PREC_FP32: return get_field_impl<float, H>(h, which, out);
    case FDTD_PREC_FP64: return get_field_impl<double, H>(h, which, out);
    case FDTD_PREC_FP16: return get_field_impl<__half, H>(h, which, out);
    default: return get_field_impl<__hip_bfloat16, H>(h, which, out);
    }
}
template <typename H>
int field_dispatch_set(fdtd_engine* h, int which, const H* in) {
    switch (h->prec) {
    case FDTD_PREC_FP32: return set_field_impl<float, H>(h, which, in);
    case FDTD_PREC_FP64: return set_field_impl<double, H>(h, which, in);
    case FDTD_PREC_FP16: return set_field_impl<__half, H>(h, which, in);
    default: return set_field_impl<__hip_bfloat16, H>(h, which, in);
    }
}

// element stored at device address -> double, on the host
double decode_elem(const fdtd_engine* h, const void* raw, double scale) {
    switch (h->prec) {
    case FDTD_PREC_FP32: return (double)*(const float*)raw;
    case FDTD_PREC_FP64: return *(const double*)raw;
    case FDTD_PREC_FP16: { __half v; memcpy(&v, raw, 2); return (double)((float)__half2float(v) * (float)scale); }
    default: { __hip_bfloat16 v; memcpy(&v, raw, 2); return (double)((float)__bfloat162float(v) * (float)scale); }
    }
}

} // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* fdtd_version(void) { return FDTD_VERSION_STR; }

int fdtd_last_error(char* buf, size_t len) {
    if (!buf || !len) return FDTD_ERR_ARG;
    snprintf(buf, len, "%s", g_last_error.c_str());
    return FDTD_OK;
}

int fdtd_device_name(char* buf, size_t len) {
    if (!buf || !len) return fail(FDTD_ERR_ARG, "null buffer");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, dev));
    snprintf(buf, len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return FDTD_OK;
}

int fdtd_create(const fdtd_grid_desc* desc, fdtd_engine** out) {
    if (!desc || !out) return fail(FDTD_ERR_ARG, "null argument");
    *out = nullptr;
    for (int n = 0; n < 3; ++n)
        if (desc->num_lines[n] < 2) return fail(FDTD_ERR_ARG, "num_lines[%d]=%u < 2", n, desc->num_lines[n]);
    if (!(desc->dT > 0)) return fail(FDTD_ERR_ARG, "dT must be > 0");
    if (desc->precision < FDTD_PREC_FP32 || desc->precision > FDTD_PREC_BF16) return fail(FDTD_ERR_ARG, "bad precision %d", desc->precision);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FDTD_ERR_NODEVICE, "no HIP device available");
    fdtd_engine* h = new fdtd_engine();
    h->prec = desc->precision;
    HIPCHK(hipGetDevice(&h->dev));
    for (int n = 0; n < 3; ++n) h->N[n] = desc->num_lines[n];
    h->dT = desc->dT;
    h->elem = h->prec == FDTD_PREC_FP64 ? 8 : (h->prec == FDTD_PREC_FP32 ? 4 : 2);
    h->acc = h->prec == FDTD_PREC_FP64 ? 8 : 4;
    const long long nzp = ((long long)h->N[2] + 15) / 16 * 16;
    h->d.nx = (int)h->N[0]; h->d.ny = (int)h->N[1]; h->d.nz = (int)h->N[2];
    h->d.sy = nzp;
    h->d.sx = nzp * h->N[1];
    h->d.sc = h->d.sx * h->N[0];
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete h; return fail(FDTD_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e)); }
    int rc;
    const size_t fbytes = (size_t)3 * h->d.sc * h->elem;
    if ((rc = h->V.alloc(fbytes)) || (rc = h->I.alloc(fbytes)) || (rc = h->step.alloc(sizeof(unsigned)))) { delete h; return rc; }
    if (hipMemsetAsync(h->V.p, 0, fbytes, h->stream) != hipSuccess || hipMemsetAsync(h->I.p, 0, fbytes, h->stream) != hipSuccess ||
        hipMemsetAsync(h->step.p, 0, sizeof(unsigned), h->stream) != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) {
        delete h;
        return fail(FDTD_ERR_HIP, "field initialisation failed");
    }
    *out = h;
    return FDTD_OK;
}

int fdtd_destroy(fdtd_engine* h) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    (void)hipStreamSynchronize(h->stream);
    delete h;
    return FDTD_OK;
}

int fdtd_set_coeffs(fdtd_engine* h, int dtype, const void* vv, const void* vi, const void* ii, const void* iv) {
    if (!h || !vv || !vi || !ii || !iv) return fail(FDTD_ERR_ARG, "null argument");
    if (h->numTS) return fail(FDTD_ERR_STATE, "coefficients must be set before the first timestep");
    int rc;
    if ((rc = upload_dense3(h, dtype, vv, h->cf_vv)) || (rc = upload_dense3(h, dtype, vi, h->cf_vi)) ||
        (rc = upload_dense3(h, dtype, ii, h->cf_ii)) || (rc = upload_dense3(h, dtype, iv, h->cf_iv)))
        return rc;
    h->cf_idx.alloc(0); h->cf_tabE.alloc(0); h->cf_tabH.alloc(0);
    h->cf_mode = CF_FULL;
    invalidate_graph(h);
    return FDTD_OK;
}

int fdtd_set_coeffs_compressed(fdtd_engine* h, const uint32_t* index, int dtype, const void* table, uint32_t n_entries) {
    if (!h || !index || !table || !n_entries) return fail(FDTD_ERR_ARG, "null argument");
    if (h->numTS) return fail(FDTD_ERR_STATE, "coefficients must be set before the first timestep");
    const size_t nd = (size_t)h->N[0] * h->N[1] * h->N[2];
    for (size_t i = 0; i < nd; ++i)
        if (index[i] >= n_entries) return fail(FDTD_ERR_RANGE, "index[%zu]=%u >= n_entries=%u", i, index[i], n_entries);
    std::vector<char> tab;
    int rc = convert_host(h, dtype, table, (size_t)n_entries * 12, tab);
    if (rc) return rc;
    // split into 8-wide E and H rows
    const size_t a = h->acc;
    std::vector<char> te((size_t)n_entries * 8 * a, 0), th((size_t)n_entries * 8 * a, 0);
    for (uint32_t r = 0; r < n_entries; ++r) {
        const char* row = tab.data() + (size_t)r * 12 * a;
        memcpy(te.data() + (size_t)r * 8 * a, row, 3 * a);            // vv
        memcpy(te.data() + ((size_t)r * 8 + 4) * a, row + 3 * a, 3 * a);  // vi
        memcpy(th.data() + (size_t)r * 8 * a, row + 6 * a, 3 * a);    // ii
        memcpy(th.data() + ((size_t)r * 8 + 4) * a, row + 9 * a, 3 * a);  // iv
    }
    const bool i16 = n_entries <= 65536;
    const size_t ib = i16 ? 2 : 4;
    if ((rc = h->cf_idx.alloc((size_t)h->d.sc * ib))) return rc;
    if ((rc = h->cf_tabE.alloc(te.size())) || (rc = h->cf_tabH.alloc(th.size()))) return rc;
    HIPCHK(hipMemsetAsync(h->cf_idx.p, 0, h->cf_idx.bytes, h->stream));
    // pack index rows into the padded layout
    std::vector<char> rows((size_t)h->N[1] * h->N[2] * ib);
    for (uint32_t x = 0; x < h->N[0]; ++x) {
        const uint32_t* src = index + (size_t)x * h->N[1] * h->N[2];
        if (i16) { uint16_t* o = (uint16_t*)rows.data(); for (size_t i = 0; i < (size_t)h->N[1] * h->N[2]; ++i) o[i] = (uint16_t)src[i]; }
        else memcpy(rows.data(), src, rows.size());
        HIPCHK(hipMemcpy2DAsync((char*)h->cf_idx.p + (size_t)x * h->d.sx * ib, (size_t)h->d.sy * ib, rows.data(), (size_t)h->N[2] * ib,
                                (size_t)h->N[2] * ib, h->N[1], hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    HIPCHK(hipMemcpyAsync(h->cf_tabE.p, te.data(), te.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->cf_tabH.p, th.data(), th.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->cf_vv.alloc(0); h->cf_vi.alloc(0); h->cf_ii.alloc(0); h->cf_iv.alloc(0);
    h->n_entries = n_entries;
    h->cf_mode = i16 ? CF_IDX16 : CF_IDX32;
    invalidate_graph(h);
    return FDTD_OK;
}

int fdtd_set_storage_scale(fdtd_engine* h, const double volt_scale[3], const double curr_scale[3]) {
    if (!h || !volt_scale || !curr_scale) return fail(FDTD_ERR_ARG, "null argument");
    if (h->numTS) return fail(FDTD_ERR_STATE, "scales must be set before the first timestep");
    for (int n = 0; n < 3; ++n) {
        if (!(volt_scale[n] > 0) || !(curr_scale[n] > 0)) return fail(FDTD_ERR_ARG, "scales must be > 0");
        if (h->prec == FDTD_PREC_FP16 || h->prec == FDTD_PREC_BF16) { h->vscale[n] = volt_scale[n]; h->cscale[n] = curr_scale[n]; }
    }
    invalidate_graph(h);
    return FDTD_OK;
}

int fdtd_set_excitation(fdtd_engine* h, const fdtd_excitation* x) {
    if (!h || !x) return fail(FDTD_ERR_ARG, "null argument");
    if (h->numTS) return fail(FDTD_ERR_STATE, "excitation must be set before the first timestep");
    if (!x->length || !x->signal_volt || !x->signal_curr) return fail(FDTD_ERR_ARG, "empty excitation signal");
    const size_t a = h->acc;
    const uint32_t nv = x->volt_count, nc = x->curr_count;
    auto check = [&](uint32_t cnt, const uint32_t* pos, const uint32_t* dir, const double* amp, const uint32_t* del) -> int {
        if (!cnt) return FDTD_OK;
        if (!pos || !dir || !amp || !del) return fail(FDTD_ERR_ARG, "null excitation list");
        for (uint32_t i = 0; i < cnt; ++i) {
            if (dir[i] > 2) return fail(FDTD_ERR_RANGE, "excitation dir %u", dir[i]);
            for (int n = 0; n < 3; ++n)
                if (pos[(size_t)n * cnt + i] >= h->N[n]) return fail(FDTD_ERR_RANGE, "excitation position out of range");
        }
        return FDTD_OK;
    };
    int rc;
    if ((rc = check(nv, x->volt_pos, x->volt_dir, x->volt_amp, x->volt_delay))) return rc;
    if ((rc = check(nc, x->curr_pos, x->curr_dir, x->curr_amp, x->curr_delay))) return rc;
    // one buffer: sigV, sigI, then per list: off(8) comp(4) delay(4) amp(a)
    const size_t L = x->length;
    size_t bytes = 2 * L * a + (size_t)(nv + nc) * (8 + 4 + 4 + a) + 64;
    if ((rc = h->exc_buf.alloc(bytes))) return rc;
    std::vector<char> hb(bytes, 0);
    size_t o = 0;
    auto put_vals = [&](const double* src, size_t n) -> size_t {
        size_t at = o;
        for (size_t i = 0; i < n; ++i) {
            if (a == 8) memcpy(hb.data() + o + i * 8, &src[i], 8);
            else { float f = (float)src[i]; memcpy(hb.data() + o + i * 4, &f, 4); }
        }
        o += n * a;
        o = (o + 7) / 8 * 8;
        return at;
    };
    const size_t o_sv = put_vals(x->signal_volt, L);
    const size_t o_si = put_vals(x->signal_curr, L);
    auto put_list = [&](uint32_t cnt, const uint32_t* pos, const uint32_t* dir, const double* amp, const uint32_t* del, ExcDev& ed, size_t o_sig) {
        ed.count = (int)cnt;
        ed.length = (int)L;
        ed.period = (int)x->period;
        ed.sig = (char*)h->exc_buf.p + o_sig;
        if (!cnt) return;
        size_t o_off = o; o += (size_t)cnt * 8;
        size_t o_comp = o; o += (size_t)cnt * 4;
        size_t o_del = o; o += (size_t)cnt * 4;
        o = (o + 7) / 8 * 8;
        std::vector<long long> keys(cnt);
        for (uint32_t i = 0; i < cnt; ++i) {
            long long off = (long long)dir[i] * h->d.sc + cell_off(h, pos[i], pos[cnt + i], pos[2 * (size_t)cnt + i]);
            memcpy(hb.data() + o_off + i * 8, &off, 8);
            int c = (int)dir[i], dl = (int)del[i];
            memcpy(hb.data() + o_comp + i * 4, &c, 4);
            memcpy(hb.data() + o_del + i * 4, &dl, 4);
            keys[i] = off;
        }
        size_t o_amp = put_vals(amp, cnt);
        std::sort(keys.begin(), keys.end());
        ed.sequential = std::adjacent_find(keys.begin(), keys.end()) != keys.end();
        ed.off = (const long long*)((char*)h->exc_buf.p + o_off);
        ed.comp = (const int*)((char*)h->exc_buf.p + o_comp);
        ed.delay = (const int*)((char*)h->exc_buf.p + o_del);
        ed.amp = (char*)h->exc_buf.p + o_amp;
    };
    // the voltage hook injects the voltage signal, the current hook the current signal
    put_list(nv, x->volt_pos, x->volt_dir, x->volt_amp, x->volt_delay, h->exc_v, o_sv);
    put_list(nc, x->curr_pos, x->curr_dir, x->curr_amp, x->curr_delay, h->exc_c, o_si);
    HIPCHK(hipMemcpyAsync(h->exc_buf.p, hb.data(), bytes, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->has_exc = true;
    invalidate_graph(h);
    return FDTD_OK;
}

int fdtd_add_upml(fdtd_engine* h, const uint32_t start[3], const uint32_t num_lines[3], int dtype,
                  const void* vv, const void* vvfo, const void* vvfn, const void* ii, const void* iifo, const void* iifn) {
    if (!h || !start || !num_lines || !vv || !vvfo || !vvfn || !ii || !iifo || !iifn) return fail(FDTD_ERR_ARG, "null argument");
    if (h->numTS) return fail(FDTD_ERR_STATE, "UPML must be added before the first timestep");
    if (h->pml_final) return fail(FDTD_ERR_STATE, "UPML slabs are fixed once stepping has been prepared");
    if (h->pml.count >= MAX_PML_SLABS) return fail(FDTD_ERR_RANGE, "at most %d UPML slabs", MAX_PML_SLABS);
    for (int n = 0; n < 3; ++n)
        if (!num_lines[n] || start[n] + num_lines[n] > h->N[n]) return fail(FDTD_ERR_RANGE, "UPML slab out of range on axis %d", n);
    for (int s = 0; s < h->pml.count; ++s) {
        const PmlSlab& b = h->pml.slab[s];
        bool overlap = true;
        for (int n = 0; n < 3; ++n)
            if (start[n] + num_lines[n] <= (uint32_t)b.s[n] || (uint32_t)(b.s[n] + b.n[n]) <= start[n]) overlap = false;
        if (overlap) return fail(FDTD_ERR_ARG, "UPML slabs overlap");
    }
    PmlSlab b{};
    for (int n = 0; n < 3; ++n) { b.s[n] = (int)start[n]; b.n[n] = (int)num_lines[n]; }
    b.za = b.s[2] / 16 * 16;
    b.nzp = (b.s[2] + b.n[2] - b.za + 15) / 16 * 16;
    b.sc = (long long)num_lines[0] * num_lines[1] * b.nzp;
    b.off = h->pml_total;
    const size_t cnt = (size_t)3 * b.sc;
    const size_t dense = (size_t)3 * num_lines[0] * num_lines[1] * num_lines[2];
    std::vector<char> all((size_t)6 * cnt * h->acc, 0);
    const void* src[6] = {vv, vvfo, vvfn, ii, iifo, iifn};
    std::vector<char> tmp;
    const size_t a = h->acc;
    for (int k = 0; k < 6; ++k) {
        int rc = convert_host(h, dtype, src[k], dense, tmp);
        if (rc) return rc;
        // dense [3][nx][ny][nz] -> padded [3][nx][ny][nzp] at z offset s[2]-za
        const size_t rows = (size_t)3 * num_lines[0] * num_lines[1];
        for (size_t r = 0; r < rows; ++r)
            memcpy(all.data() + ((size_t)k * cnt + r * b.nzp + (b.s[2] - b.za)) * a, tmp.data() + r * num_lines[2] * a, num_lines[2] * a);
    }
    // offsets are in elements per array; each array covers all slabs, comp-major per slab
    h->pml.slab[h->pml.count++] = b;
    h->pml_total += (long long)cnt;
    h->pml_host_coef.push_back(std::move(all));
    h->pml_dirty = true;
    invalidate_graph(h);
    return FDTD_OK;
}

int fdtd_add_ade(fdtd_engine* h, uint32_t count, const uint32_t* pos, int dtype, const void* v_int, const void* v_ext,
                 const void* v_lor, const void* i_int, const void* i_ext, const void* i_lor, uint32_t flags) {
    if (!h || !pos) return fail(FDTD_ERR_ARG, "null argument");
    if (h->numTS) return fail(FDTD_ERR_STATE, "ADE must be added before the first timestep");
    if (!count) return FDTD_OK;
    if ((flags & 1) && (!v_int || !v_ext)) return fail(FDTD_ERR_ARG, "voltage ADE lists missing");
    if ((flags & 2) && (!i_int || !i_ext)) return fail(FDTD_ERR_ARG, "current ADE lists missing");
    if ((flags & 4) && !v_lor) return fail(FDTD_ERR_ARG, "voltage Lorentz list missing");
    if ((flags & 8) && !i_lor) return fail(FDTD_ERR_ARG, "current Lorentz list missing");
    std::vector<long long> off(count);
    for (uint32_t i = 0; i < count; ++i) {
        for (int n = 0; n < 3; ++n)
            if (pos[(size_t)n * count + i] >= h->N[n]) return fail(FDTD_ERR_RANGE, "ADE position out of range");
        off[i] = cell_off(h, pos[i], pos[count + i], pos[2 * (size_t)count + i]);
    }
    AdeHost* a = new AdeHost();
    a->count = count;
    a->flags = flags;
    const size_t n3 = (size_t)3 * count, ab = h->acc;
    int rc;
    if ((rc = a->off.alloc(count * sizeof(long long))) || (rc = a->coef.alloc(6 * n3 * ab)) || (rc = a->state.alloc(4 * n3 * ab))) { delete a; return rc; }
    std::vector<char> all(6 * n3 * ab, 0), tmp;
    const void* lists[6] = {v_int, v_ext, v_lor, i_int, i_ext, i_lor};
    for (int k = 0; k < 6; ++k) {
        if (!lists[k]) continue;
        if ((rc = convert_host(h, dtype, lists[k], n3, tmp))) { delete a; return rc; }
        memcpy(all.data() + k * n3 * ab, tmp.data(), n3 * ab);
    }
    if (hipMemcpyAsync(a->off.p, off.data(), count * sizeof(long long), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipMemcpyAsync(a->coef.p, all.data(), all.size(), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipMemsetAsync(a->state.p, 0, a->state.bytes, h->stream) != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) {
        delete a;
        return fail(FDTD_ERR_HIP, "ADE upload failed");
    }
    h->ade.push_back(a);
    invalidate_graph(h);
    return FDTD_OK;
}

int fdtd_set_tfsf(fdtd_engine* h, int which, uint32_t count, const uint32_t* pos, const uint32_t* comp, const uint32_t* delay,
                  const double* delta, const double* amp, uint32_t max_delay) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    if (h->numTS) return fail(FDTD_ERR_STATE, "TFSF must be set before the first timestep");
    if (!h->has_exc) return fail(FDTD_ERR_STATE, "set the excitation before the TFSF tables");
    if (which != FDTD_VOLT && which != FDTD_CURR) return fail(FDTD_ERR_ARG, "bad field selector");
    TfsfDev& t = which == FDTD_VOLT ? h->tfsf_v : h->tfsf_c;
    DevBuf& buf = which == FDTD_VOLT ? h->tfsf_buf_v : h->tfsf_buf_c;
    t = TfsfDev{};
    if (!count) return FDTD_OK;
    if (!pos || !comp || !delay || !delta || !amp) return fail(FDTD_ERR_ARG, "null TFSF list");
    const size_t a = h->acc;
    size_t bytes = (size_t)count * (8 + 4 + 4 + 2 * a) + 64;
    int rc = buf.alloc(bytes);
    if (rc) return rc;
    std::vector<char> hb(bytes, 0);
    size_t o_off = 0, o_comp = (size_t)count * 8, o_del = o_comp + (size_t)count * 4;
    size_t o_dd = (o_del + (size_t)count * 4 + 7) / 8 * 8, o_amp = o_dd + (size_t)count * a;
    for (uint32_t i = 0; i < count; ++i) {
        for (int n = 0; n < 3; ++n)
            if (pos[(size_t)n * count + i] >= h->N[n]) return fail(FDTD_ERR_RANGE, "TFSF position out of range");
        if (comp[i] > 2) return fail(FDTD_ERR_RANGE, "TFSF component");
        if (delay[i] + 1 > max_delay + 1) return fail(FDTD_ERR_RANGE, "TFSF delay beyond max_delay");
        long long off = (long long)comp[i] * h->d.sc + cell_off(h, pos[i], pos[count + i], pos[2 * (size_t)count + i]);
        int c = (int)comp[i], dl = (int)delay[i];
        memcpy(hb.data() + o_off + i * 8, &off, 8);
        memcpy(hb.data() + o_comp + i * 4, &c, 4);
        memcpy(hb.data() + o_del + i * 4, &dl, 4);
        if (a == 8) { memcpy(hb.data() + o_dd + i * 8, &delta[i], 8); memcpy(hb.data() + o_amp + i * 8, &amp[i], 8); }
        else { float f = (float)delta[i], g = (float)amp[i]; memcpy(hb.data() + o_dd + i * 4, &f, 4); memcpy(hb.data() + o_amp + i * 4, &g, 4); }
    }
    HIPCHK(hipMemcpyAsync(buf.p, hb.data(), bytes, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    char* b = (char*)buf.p;
    t.count = (int)count;
    t.length = h->exc_v.length;
    t.period = h->exc_v.period;
    t.max_delay = (int)max_delay;
    t.off = (const long long*)(b + o_off);
    t.comp = (const int*)(b + o_comp);
    t.delay = (const int*)(b + o_del);
    t.delta = b + o_dd;
    t.amp = b + o_amp;
    // voltage injection uses the current signal and vice versa (engine_ext_tfsf.cpp:54,145)
    t.sig = which == FDTD_VOLT ? h->exc_c.sig : h->exc_v.sig;
    invalidate_graph(h);
    return FDTD_OK;
}

int fdtd_add_probe(fdtd_engine* h, int kind, const uint32_t start[3], const uint32_t stop[3], int norm_dir, uint32_t inside_mask,
                   double weight, uint32_t td_interval, uint32_t fd_interval, const double* freqs, uint32_t n_freq, int* id) {
    if (!h || !start || !stop || !id) return fail(FDTD_ERR_ARG, "null argument");
    if (h->numTS || h->sampled0) return fail(FDTD_ERR_STATE, "probes must be added before the first timestep");
    if (n_freq && (!freqs || !fd_interval)) return fail(FDTD_ERR_ARG, "FD probe needs frequencies and an interval");
    for (int n = 0; n < 3; ++n)
        if (start[n] >= h->N[n] || stop[n] >= h->N[n]) return fail(FDTD_ERR_RANGE, "probe out of range");
    ProbeHost* p = new ProbeHost();
    p->kind = kind;
    p->norm_dir = norm_dir;
    p->weight = weight;
    p->td_interval = td_interval;
    p->fd_interval = fd_interval;
    p->freqs.assign(freqs, freqs + n_freq);
    auto add = [&](int n, uint32_t x, uint32_t y, uint32_t z, int sgn) {
        p->off.push_back((long long)n * h->d.sc + cell_off(h, x, y, z));
        p->comp.push_back(n);
        p->sign.push_back((signed char)sgn);
    };
    if (kind == FDTD_PROBE_VOLTAGE) {
        // Engine_Interface_FDTD::CalcVoltageIntegral, FDTD/engine_interface_fdtd.cpp:200-226
        int ndiff = (start[0] != stop[0]) + (start[1] != stop[1]) + (start[2] != stop[2]);
        if (ndiff != 1) { delete p; return fail(FDTD_ERR_ARG, "voltage probe must be a line along one axis"); }
        for (int n = 0; n < 3; ++n) {
            if (start[n] < stop[n]) {
                uint32_t pos[3] = {start[0], start[1], start[2]};
                for (; pos[n] < stop[n]; ++pos[n]) add(n, pos[0], pos[1], pos[2], +1);
            } else if (start[n] > stop[n]) {
                uint32_t pos[3] = {stop[0], stop[1], stop[2]};
                for (; pos[n] < start[n]; ++pos[n]) add(n, pos[0], pos[1], pos[2], -1);
            }
        }
    } else if (kind == FDTD_PROBE_CURRENT) {
        // ProcessCurrent::CalcIntegral, Common/processcurrent.cpp:93-168
        if (norm_dir < 0 || norm_dir > 2 || start[norm_dir] != stop[norm_dir]) { delete p; return fail(FDTD_ERR_ARG, "current probe needs a valid normal"); }
        const bool si[3] = {(inside_mask & 1) != 0, (inside_mask & 2) != 0, (inside_mask & 4) != 0};
        const bool so[3] = {(inside_mask & 8) != 0, (inside_mask & 16) != 0, (inside_mask & 32) != 0};
        const uint32_t* a = start; const uint32_t* b = stop;
        switch (norm_dir) {
        case 0:
            if (so[0] && si[2]) for (uint32_t i = a[1] + 1; i <= b[1]; ++i) add(1, b[0], i, a[2], +1);
            if (so[0] && so[1]) for (uint32_t i = a[2] + 1; i <= b[2]; ++i) add(2, b[0], b[1], i, +1);
            if (si[0] && so[2]) for (uint32_t i = a[1] + 1; i <= b[1]; ++i) add(1, a[0], i, b[2], -1);
            if (si[0] && si[1]) for (uint32_t i = a[2] + 1; i <= b[2]; ++i) add(2, a[0], a[1], i, -1);
            break;
        case 1:
            if (si[0] && si[1]) for (uint32_t i = a[2] + 1; i <= b[2]; ++i) add(2, a[0], a[1], i, +1);
            if (so[1] && so[2]) for (uint32_t i = a[0] + 1; i <= b[0]; ++i) add(0, i, b[1], b[2], +1);
            if (so[0] && so[1]) for (uint32_t i = a[2] + 1; i <= b[2]; ++i) add(2, b[0], b[1], i, -1);
            if (si[1] && si[2]) for (uint32_t i = a[0] + 1; i <= b[0]; ++i) add(0, i, a[1], a[2], -1);
            break;
        default:
            if (si[1] && si[2]) for (uint32_t i = a[0] + 1; i <= b[0]; ++i) add(0, i, a[1], a[2], +1);
            if (so[0] && si[2]) for (uint32_t i = a[1] + 1; i <= b[1]; ++i) add(1, b[0], i, a[2], +1);
            if (so[1] && so[2]) for (uint32_t i = a[0] + 1; i <= b[0]; ++i) add(0, i, b[1], b[2], -1);
            if (si[0] && so[2]) for (uint32_t i = a[1] + 1; i <= b[1]; ++i) add(1, a[0], i, b[2], -1);
            break;
        }
    } else {
        delete p;
        return fail(FDTD_ERR_ARG, "unknown probe kind %d", kind);
    }
    int rc;
    if (n_freq) {
        if ((rc = p->fd.alloc((size_t)n_freq * 2 * sizeof(double))) || (rc = p->dfreq.alloc((size_t)n_freq * sizeof(double)))) { delete p; return rc; }
        if (hipMemsetAsync(p->fd.p, 0, p->fd.bytes, h->stream) != hipSuccess ||
            hipMemcpyAsync(p->dfreq.p, freqs, (size_t)n_freq * sizeof(double), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess) { delete p; return fail(FDTD_ERR_HIP, "probe upload failed"); }
    }
    *id = (int)h->probes.size();
    h->probes.push_back(p);
    h->probes_dirty = true;
    return FDTD_OK;
}

int fdtd_probe_num_samples(fdtd_engine* h, int id, uint32_t* n) {
    if (!h || !n) return fail(FDTD_ERR_ARG, "null argument");
    if (id < 0 || id >= (int)h->probes.size()) return fail(FDTD_ERR_RANGE, "bad probe id");
    ProbeHost* p = h->probes[id];
    *n = (p->td_interval && h->sampled0) ? h->numTS / p->td_interval + 1 : 0;
    return FDTD_OK;
}

int fdtd_read_probe(fdtd_engine* h, int id, double* t, double* val, uint32_t max_n) {
    if (!h || !t || !val) return fail(FDTD_ERR_ARG, "null argument");
    uint32_t n;
    int rc = fdtd_probe_num_samples(h, id, &n);
    if (rc) return rc;
    n = std::min(n, max_n);
    if (!n) return FDTD_OK;
    std::vector<double> buf((size_t)2 * n);
    HIPCHK(hipMemcpyAsync(buf.data(), h->probes[id]->td.p, buf.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (uint32_t i = 0; i < n; ++i) { t[i] = buf[2 * i]; val[i] = buf[2 * i + 1]; }
    return FDTD_OK;
}

int fdtd_read_probe_fd(fdtd_engine* h, int id, double* re, double* im) {
    if (!h || !re || !im) return fail(FDTD_ERR_ARG, "null argument");
    if (id < 0 || id >= (int)h->probes.size()) return fail(FDTD_ERR_RANGE, "bad probe id");
    ProbeHost* p = h->probes[id];
    const size_t nf = p->freqs.size();
    if (!nf) return FDTD_OK;
    std::vector<double> buf(2 * nf);
    HIPCHK(hipMemcpyAsync(buf.data(), p->fd.p, buf.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (size_t k = 0; k < nf; ++k) { re[k] = buf[2 * k]; im[k] = buf[2 * k + 1]; }
    return FDTD_OK;
}

int fdtd_set_tuning(fdtd_engine* h, int variant, int tile_y, int chunk_x) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    if (variant < 0 || variant > 3) return fail(FDTD_ERR_ARG, "unknown kernel variant %d", variant);
    if (tile_y != 1 && tile_y != 2 && tile_y != 4) return fail(FDTD_ERR_ARG, "tile_y must be 1, 2 or 4");
    if (chunk_x < 1) return fail(FDTD_ERR_ARG, "chunk_x must be >= 1");
    h->variant = variant;
    h->tile_y = tile_y;
    h->chunk_x = chunk_x;
    invalidate_graph(h);
    return FDTD_OK;
}

int fdtd_set_use_graph(fdtd_engine* h, int enable) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    h->use_graph = enable != 0;
    return FDTD_OK;
}

int fdtd_set_kernel_timing(fdtd_engine* h, int enable) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    h->timing = enable != 0;
    h->t_volt = h->t_curr = 0;
    h->n_volt = h->n_curr = 0;
    return FDTD_OK;
}

int fdtd_get_kernel_times(fdtd_engine* h, double* volt_ms, double* curr_ms, uint32_t* nv, uint32_t* nc) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    if (volt_ms) *volt_ms = h->t_volt;
    if (curr_ms) *curr_ms = h->t_curr;
    if (nv) *nv = h->n_volt;
    if (nc) *nc = h->n_curr;
    return FDTD_OK;
}

int fdtd_iterate(fdtd_engine* h, uint32_t nsteps) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    int rc = ready_check(h);
    if (rc) return rc;
    if ((rc = finalize_pml(h))) return rc;
    if ((rc = finalize_probes(h, h->numTS + nsteps))) return rc;
    if (!h->sampled0) {
        // ProcessingArray::Process before the loop (openems.cpp:1277-1278): sample at numTS=0
        switch (h->prec) {
        case FDTD_PREC_FP32: rc = launch_step_end<float>(h, 0); break;
        case FDTD_PREC_FP64: rc = launch_step_end<double>(h, 0); break;
        case FDTD_PREC_FP16: rc = launch_step_end<__half>(h, 0); break;
        default: rc = launch_step_end<__hip_bfloat16>(h, 0); break;
        }
        if (rc) return rc;
        h->sampled0 = true;
    }
    if (!nsteps) return FDTD_OK;
    if (h->timing) {
        // events for every launch of the two main kernels, harvested after the loop
        const size_t need = (size_t)4 * nsteps;
        while (h->ev.size() < need) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            h->ev.push_back(e);
        }
        h->ev_used = 0;
        for (uint32_t s = 0; s < nsteps; ++s)
            if ((rc = launch_step(h, true))) return rc;
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventSynchronize(h->ev[need - 1]));
        for (uint32_t s = 0; s < nsteps; ++s) {
            float a = 0, b = 0;
            HIPCHK(hipEventElapsedTime(&a, h->ev[4 * s], h->ev[4 * s + 1]));
            HIPCHK(hipEventElapsedTime(&b, h->ev[4 * s + 2], h->ev[4 * s + 3]));
            h->t_volt += a; h->t_curr += b;
        }
        h->n_volt += nsteps; h->n_curr += nsteps;
        h->numTS += nsteps;
        return FDTD_OK;
    }
    if (!h->use_graph) {
        for (uint32_t s = 0; s < nsteps; ++s)
            if ((rc = launch_step(h, false))) return rc;
        HIPCHK(hipGetLastError());
        h->numTS += nsteps;
        return FDTD_OK;
    }
    const int G = 16;
    uint32_t left = nsteps;
    if (left >= (uint32_t)G) {
        if (!h->gexec) { if ((rc = capture_graph(h, G, &h->graph, &h->gexec))) return rc; }
        while (left >= (uint32_t)G) { HIPCHK(hipGraphLaunch(h->gexec, h->stream)); left -= G; }
    }
    if (left) {
        if (!h->gexec1) { if ((rc = capture_graph(h, 1, &h->graph1, &h->gexec1))) return rc; }
        while (left) { HIPCHK(hipGraphLaunch(h->gexec1, h->stream)); --left; }
    }
    h->numTS += nsteps;
    return FDTD_OK;
}

int fdtd_synchronize(fdtd_engine* h) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipGetLastError());
    return FDTD_OK;
}

int fdtd_get_num_timesteps(fdtd_engine* h, uint32_t* ts) {
    if (!h || !ts) return fail(FDTD_ERR_ARG, "null argument");
    *ts = h->numTS;
    return FDTD_OK;
}

int fdtd_calc_energy(fdtd_engine* h, double* energy) {
    if (!h || !energy) return fail(FDTD_ERR_ARG, "null argument");
    switch (h->prec) {
    case FDTD_PREC_FP32: return energy_impl<float>(h, energy);
    case FDTD_PREC_FP64: return energy_impl<double>(h, energy);
    case FDTD_PREC_FP16: return energy_impl<__half>(h, energy);
    default: return energy_impl<__hip_bfloat16>(h, energy);
    }
}

int fdtd_get_field(fdtd_engine* h, int which, int dtype, void* out) {
    if (!h || !out) return fail(FDTD_ERR_ARG, "null argument");
    if (which != FDTD_VOLT && which != FDTD_CURR) return fail(FDTD_ERR_ARG, "bad field selector");
    if (dtype == FDTD_DTYPE_F32) return field_dispatch_get<float>(h, which, (float*)out);
    if (dtype == FDTD_DTYPE_F64) return field_dispatch_get<double>(h, which, (double*)out);
    return fail(FDTD_ERR_ARG, "bad dtype");
}

int fdtd_set_field(fdtd_engine* h, int which, int dtype, const void* in) {
    if (!h || !in) return fail(FDTD_ERR_ARG, "null argument");
    if (which != FDTD_VOLT && which != FDTD_CURR) return fail(FDTD_ERR_ARG, "bad field selector");
    if (dtype == FDTD_DTYPE_F32) return field_dispatch_set<float>(h, which, (const float*)in);
    if (dtype == FDTD_DTYPE_F64) return field_dispatch_set<double>(h, which, (const double*)in);
    return fail(FDTD_ERR_ARG, "bad dtype");
}

int fdtd_get_value(fdtd_engine* h, int which, uint32_t n, uint32_t x, uint32_t y, uint32_t z, double* v) {
    if (!h || !v) return fail(FDTD_ERR_ARG, "null argument");
    if (which != FDTD_VOLT && which != FDTD_CURR) return fail(FDTD_ERR_ARG, "bad field selector");
    if (n > 2 || x >= h->N[0] || y >= h->N[1] || z >= h->N[2]) return fail(FDTD_ERR_RANGE, "position out of range");
    const long long off = (long long)n * h->d.sc + cell_off(h, x, y, z);
    char raw[8];
    const DevBuf& F = which == FDTD_VOLT ? h->V : h->I;
    HIPCHK(hipMemcpyAsync(raw, (char*)F.p + off * h->elem, h->elem, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    *v = decode_elem(h, raw, which == FDTD_VOLT ? h->vscale[n] : h->cscale[n]);
    return FDTD_OK;
}

int fdtd_set_value(fdtd_engine* h, int which, uint32_t n, uint32_t x, uint32_t y, uint32_t z, double v) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    if (which != FDTD_VOLT && which != FDTD_CURR) return fail(FDTD_ERR_ARG, "bad field selector");
    if (n > 2 || x >= h->N[0] || y >= h->N[1] || z >= h->N[2]) return fail(FDTD_ERR_RANGE, "position out of range");
    const long long off = (long long)n * h->d.sc + cell_off(h, x, y, z);
    char raw[8];
    const double s = which == FDTD_VOLT ? h->vscale[n] : h->cscale[n];
    switch (h->prec) {
    case FDTD_PREC_FP32: { float f = (float)v; memcpy(raw, &f, 4); break; }
    case FDTD_PREC_FP64: memcpy(raw, &v, 8); break;
    case FDTD_PREC_FP16: { __half q = __float2half_rn((float)v * (float)(1.0 / s)); memcpy(raw, &q, 2); break; }
    default: { __hip_bfloat16 q = __float2bfloat16((float)v * (float)(1.0 / s)); memcpy(raw, &q, 2); break; }
    }
    DevBuf& F = which == FDTD_VOLT ? h->V : h->I;
    HIPCHK(hipMemcpyAsync((char*)F.p + off * h->elem, raw, h->elem, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return FDTD_OK;
}

int fdtd_copy_bench(uint64_t bytes, int reps, double* gbps) {
    if (!gbps || reps < 1 || bytes < 4096) return fail(FDTD_ERR_ARG, "bad argument");
    const long long n = (long long)(bytes / 16);
    DevBuf a, b;
    int rc;
    if ((rc = a.alloc((size_t)n * 16)) || (rc = b.alloc((size_t)n * 16))) return rc;
    hipStream_t st;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipMemsetAsync(a.p, 0, a.bytes, st));
    const int blocks = 256 * 16;
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, st, a.as<float4>(), b.as<float4>(), n);   // warm-up
    HIPCHK(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, st, a.as<float4>(), b.as<float4>(), n);
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    *gbps = 2.0 * (double)n * 16 * reps / (ms * 1e-3) / 1e9;
    return FDTD_OK;
}

int fdtd_device_bytes(fdtd_engine* h, uint64_t* bytes) {
    if (!h || !bytes) return fail(FDTD_ERR_ARG, "null argument");
    uint64_t b = h->V.bytes + h->I.bytes + h->cf_vv.bytes + h->cf_vi.bytes + h->cf_ii.bytes + h->cf_iv.bytes + h->cf_idx.bytes +
                 h->cf_tabE.bytes + h->cf_tabH.bytes + h->pml_coef.bytes + h->pml_flux.bytes + h->exc_buf.bytes;
    for (auto a : h->ade) b += a->off.bytes + a->coef.bytes + a->state.bytes;
    *bytes = b;
    return FDTD_OK;
}

} // extern "C"
