// fdtd_device.h -- device side of the MI355X FDTD engine: storage traits,
// kernel argument blocks and the update / hook / probe kernels.  Included by
// fdtd_engine.hip (single domains) and fdtd_batch.hip (batched domains).
#pragma once
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
// start at za = s[2] rounded down to a multiple of the main arrays' vector
// width (16 bytes of storage) and padded to nzp, a multiple of it, so a
// 16-byte-aligned vector of consecutive z cells of the main arrays maps to
// an aligned vector of the slab arrays.  Padding cells hold zeros.
struct PmlSlab {
    int s[3];              // first line per axis
    int n[3];              // lines per axis
    int za;                // aligned first z of the slab arrays
    int nzp;               // padded z extent of the slab arrays
    long long off;         // element offset of this slab's component 0
    long long sc;          // component stride inside the slab (nx*ny*nzp)
    long long ioff;        // cell offset of this slab in the coefficient index arrays
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
    // compressed form used by the vector kernels when present: per slab cell
    // a uint16 row index, rows of 12 {k0,k1,k2,_, fo0,fo1,fo2,_, fn0,fn1,fn2,_}
    const unsigned short* vidx; const A* vtab;
    const unsigned short* cidx; const A* ctab;
};

// coefficient access modes; CF_FACT16 (float arithmetic only): a uint16
// index into rows {a0,a1,a2,_,m0,m1,m2,_} and b = m * ((gx[x] * gy[y]) * gz[z])
// from per-line mesh factors
enum { CF_FULL = 0, CF_IDX16 = 1, CF_IDX32 = 2, CF_FACT16 = 3 };

template <typename A>
struct CoeffPtrs {
    const A* a;            // full: vv (E) or ii (H) [3][cells]; table: [U][8] rows {a0,a1,a2,_,b0,b1,b2,_}
    const A* b;            // full: vi (E) or iv (H)
    const void* idx;       // per-cell index (compressed modes)
    // CF_FACT16 mesh factors: component n's x factors at g + n*gs, y factors
    // at + gy, z factors at + gz (16-aligned, z padded like the fields)
    const A* g;
    int gy, gz, gs;
};

// cells z0..z0+vz-1 at (x,y) that lie in slab b: mask bit k = cell z0+k;
// base = slab-array offset of z0 (aligned like z0).  Returns the mask.
__device__ __forceinline__ unsigned slab_lanes(const PmlSlab& b, int x, int y, int z0, int vz, long long& base) {
    const int lx = x - b.s[0], ly = y - b.s[1];
    if ((unsigned)lx >= (unsigned)b.n[0] || (unsigned)ly >= (unsigned)b.n[1]) return 0u;
    unsigned m = 0;
    for (int k = 0; k < vz; ++k)
        if ((unsigned)(z0 + k - b.s[2]) < (unsigned)b.n[2]) m |= 1u << k;
    base = b.off + ((long long)lx * b.n[1] + ly) * b.nzp + (z0 - b.za);
    return m;
}

// ---------------------------------------------------------------------------
// Main updates (Engine_SSE_Compressed::UpdateVoltages/UpdateCurrents,
// FDTD/engine_sse_compressed.cpp:48-263, in the operation order of
// FDTD/engine.cpp:108-210; the z-1 neighbour at z=0 is a zero-filled lane,
// :95-153) with UPML Pre/Post fused per cell (engine_ext_upml.cpp:57-234):
//   pre : h = pv*V - fo*F ; V = F ; F = h
//   main: V = V*vv + vi*curl
//   post: h = F ; F = V ; V = h + fn*F
// Vectorised: each lane updates VZ consecutive z cells with one wide
// load/store per field component (16 bytes for FP32, FP16/BF16 and FP64), so
// the number of memory instructions per cell drops VZ-fold; the z-1 (E) and
// z+1 (H) neighbours come from the neighbouring element in the same vector,
// and only the first/last element needs one scalar load.  Per-cell arithmetic
// is the same as volt_cell/curr_cell.
// ---------------------------------------------------------------------------
template <typename S> struct VecWidth { enum { value = 16 / sizeof(S) }; };

template <typename S, int VZ>
struct alignas(sizeof(S) * VZ) SVec {
    S v[VZ];
};

// Streaming (nontemporal) access for data one launch touches once: the
// updated field's own components, the coefficient arrays and index, UPML flux
// and slab coefficients.  They bypass the caches' normal retention, so the
// other field's planes, which the x-1 and y-1 neighbours read again, stay in
// L2.  Loads and stores of 4, 8, 16 or 32 bytes.
template <int B> struct RawBits;
template <> struct RawBits<4> { typedef unsigned int T; };
template <> struct RawBits<8> { typedef unsigned int T __attribute__((ext_vector_type(2))); };
template <> struct RawBits<16> { typedef unsigned int T __attribute__((ext_vector_type(4))); };
template <> struct RawBits<32> { typedef unsigned int T __attribute__((ext_vector_type(8))); };

template <typename V>
__device__ __forceinline__ V ld_stream(const V* p) {
    typedef typename RawBits<sizeof(V)>::T R;
    const R r = __builtin_nontemporal_load(reinterpret_cast<const R*>(p));
    V v;
    __builtin_memcpy(&v, &r, sizeof(V));
    return v;
}

template <typename V>
__device__ __forceinline__ void st_stream(V* p, const V& v) {
    typedef typename RawBits<sizeof(V)>::T R;
    R r;
    __builtin_memcpy(&r, &v, sizeof(V));
    __builtin_nontemporal_store(r, reinterpret_cast<R*>(p));
}

template <typename S, int VZ>
__device__ __forceinline__ void vload(const S* p, typename StoreTraits<S>::A* out, typename StoreTraits<S>::A s) {
    const SVec<S, VZ> t = *reinterpret_cast<const SVec<S, VZ>*>(p);
#pragma unroll
    for (int k = 0; k < VZ; ++k) out[k] = StoreTraits<S>::load(&t.v[k], s);
}

template <typename S, int VZ>
__device__ __forceinline__ void vload_stream(const S* p, typename StoreTraits<S>::A* out, typename StoreTraits<S>::A s) {
    const SVec<S, VZ> t = ld_stream(reinterpret_cast<const SVec<S, VZ>*>(p));
#pragma unroll
    for (int k = 0; k < VZ; ++k) out[k] = StoreTraits<S>::load(&t.v[k], s);
}

template <typename S, int VZ>
__device__ __forceinline__ void vstore(S* p, const typename StoreTraits<S>::A* in, typename StoreTraits<S>::A inv_s, int valid) {
    SVec<S, VZ> t;
#pragma unroll
    for (int k = 0; k < VZ; ++k) StoreTraits<S>::store(&t.v[k], in[k], inv_s);
    if (valid >= VZ) {
        st_stream(reinterpret_cast<SVec<S, VZ>*>(p), t);
    } else {
#pragma unroll
        for (int k = 0; k < VZ; ++k)
            if (k < valid) p[k] = t.v[k];
    }
}

// UPML coefficients of one vector of slab cells (component n): from the
// compressed rows when present, else from the full slab arrays
template <typename A, int VZ, bool ONE>
__device__ __forceinline__ void pml_coeffs(const A* full0, const A* full1, const A* full2, long long q,
                                           const A* tab, const unsigned short* rows, int n, A k0[VZ], A kfo[VZ], A kfn[VZ]) {
    if (tab) {
        bool same = ONE;
#pragma unroll
        for (int k = 1; k < VZ; ++k) same &= rows[k] == rows[0];
        if (same) {          // one row for the whole vector: the common case
            const A* r = tab + (size_t)rows[0] * 12;
            const A a = r[n], b = r[4 + n], c = r[8 + n];
#pragma unroll
            for (int k = 0; k < VZ; ++k) { k0[k] = a; kfo[k] = b; kfn[k] = c; }
        } else {
#pragma unroll
            for (int k = 0; k < VZ; ++k) {
                const A* r = tab + (size_t)rows[k] * 12;
                k0[k] = r[n]; kfo[k] = r[4 + n]; kfn[k] = r[8 + n];
            }
        }
    } else {
        const SVec<A, VZ> a = ld_stream(reinterpret_cast<const SVec<A, VZ>*>(full0 + q));
        const SVec<A, VZ> b = ld_stream(reinterpret_cast<const SVec<A, VZ>*>(full1 + q));
        const SVec<A, VZ> c = ld_stream(reinterpret_cast<const SVec<A, VZ>*>(full2 + q));
#pragma unroll
        for (int k = 0; k < VZ; ++k) { k0[k] = a.v[k]; kfo[k] = b.v[k]; kfn[k] = c.v[k]; }
    }
}

// UPML form of one component for the lanes pm of a vector in one slab
// (Engine_Ext_UPML pre/post around the main update, fused): per lane
// h = k*V - fo*F; F = F*a + b*curl; V = h + fn*F.  q = slab-array offset.
template <typename A, int VZ, bool ONE>
__device__ __forceinline__ void upml_lanes(A* flux, const A* k0, const A* kfo, const A* kfn, const A* tab,
                                           const unsigned short* rows, long long q, int n, unsigned pm, const A* v,
                                           const A* a, const A* b, const A* curl, A* out) {
    A F[VZ], kv[VZ], fo[VZ], fn[VZ];
    vload_stream<A, VZ>(flux + q, F, A(1));
    pml_coeffs<A, VZ, ONE>(k0, kfo, kfn, q, tab, rows, n, kv, fo, fn);
#pragma unroll
    for (int k = 0; k < VZ; ++k) {
        if (pm & (1u << k)) {
            A h = kv[k] * v[k] - fo[k] * F[k];
            A t = F[k];
            t = t * a[k];
            t = t + b[k] * curl[k];
            F[k] = t;
            out[k] = h + fn[k] * t;
        }
    }
    vstore<A, VZ>(flux + q, F, A(1), VZ);
}

// the (at most two, see fdtd_engine's slab check) slabs holding cells of the
// vector z0..z0+VZ-1 at (x, y), with their lane masks and slab-array offsets
struct SlabHits {
    int s[2];
    unsigned m[2];
    long long b[2];
};
template <int VZ>
__device__ __forceinline__ SlabHits slab_hits(const PmlSet& ps, int x, int y, int z0, unsigned lanes) {
    SlabHits h{{-1, -1}, {0u, 0u}, {0, 0}};
    int c = 0;
    for (int s = 0; s < ps.count && c < 2; ++s) {
        long long pb = 0;
        const unsigned m = slab_lanes(ps.slab[s], x, y, z0, VZ, pb) & lanes;
        if (!m) continue;
        h.s[c] = s; h.m[c] = m; h.b[c] = pb;
        ++c;
    }
    return h;
}

// Table reads of one row per vector where all its cells share it: for FP32
// and FP64 storage (large dipole FP32 30.6 -> 29.8 ms per step, dispersive
// 0.894 -> 0.879 ms); 2-byte storage measured 1-2% slower with the second
// path (registers) and gathers per cell.
template <typename S> struct OneRow { enum { value = sizeof(S) >= 4 }; };

// rows {a0,a1,a2,_,b0,b1,b2,_} of the cells' table entries c[0..VZ-1]; with
// ONE, a vector whose cells share one entry (runs of one material along z)
// loads it once
template <typename A, int VZ, bool ONE>
__device__ __forceinline__ void table_rows(const A* tab, const unsigned c[VZ], A a[3][VZ], A b[3][VZ]) {
    bool same = ONE;
#pragma unroll
    for (int k = 1; k < VZ; ++k) same &= c[k] == c[0];
    if (same) {
        const A* row = tab + (size_t)c[0] * 8;
        const SVec<A, 4> r0 = *reinterpret_cast<const SVec<A, 4>*>(row);
        const SVec<A, 4> r1 = *reinterpret_cast<const SVec<A, 4>*>(row + 4);
#pragma unroll
        for (int k = 0; k < VZ; ++k)
#pragma unroll
            for (int n = 0; n < 3; ++n) { a[n][k] = r0.v[n]; b[n][k] = r1.v[n]; }
    } else {
#pragma unroll
        for (int k = 0; k < VZ; ++k) {
            const A* row = tab + (size_t)c[k] * 8;
            const SVec<A, 4> r0 = *reinterpret_cast<const SVec<A, 4>*>(row);
            const SVec<A, 4> r1 = *reinterpret_cast<const SVec<A, 4>*>(row + 4);
#pragma unroll
            for (int n = 0; n < 3; ++n) { a[n][k] = r0.v[n]; b[n][k] = r1.v[n]; }
        }
    }
}

template <typename A, int CM, int VZ, bool ONE>
__device__ __forceinline__ void vload_coeffs(const CoeffPtrs<A>& cf, long long p, long long sc, int x, int y, int z0,
                                             A a[3][VZ], A b[3][VZ]) {
    if (CM == CF_FACT16) {
        const SVec<unsigned short, VZ> ti = ld_stream(reinterpret_cast<const SVec<unsigned short, VZ>*>((const unsigned short*)cf.idx + p));
        unsigned c[VZ];
#pragma unroll
        for (int k = 0; k < VZ; ++k) c[k] = ti.v[k];
        table_rows<A, VZ, ONE>(cf.a, c, a, b);
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            const A* g = cf.g + n * cf.gs;
            const A gxy = g[x] * g[cf.gy + y];
            const SVec<A, VZ> gz = *reinterpret_cast<const SVec<A, VZ>*>(g + cf.gz + z0);
#pragma unroll
            for (int k = 0; k < VZ; ++k) b[n][k] = b[n][k] * (gxy * gz.v[k]);
        }
    } else if (CM == CF_FULL) {
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            const SVec<A, VZ> ta = ld_stream(reinterpret_cast<const SVec<A, VZ>*>(cf.a + n * sc + p));
            const SVec<A, VZ> tb = ld_stream(reinterpret_cast<const SVec<A, VZ>*>(cf.b + n * sc + p));
#pragma unroll
            for (int k = 0; k < VZ; ++k) { a[n][k] = ta.v[k]; b[n][k] = tb.v[k]; }
        }
    } else {
        unsigned c[VZ];
        if (CM == CF_IDX16) {
            const SVec<unsigned short, VZ> ti = ld_stream(reinterpret_cast<const SVec<unsigned short, VZ>*>((const unsigned short*)cf.idx + p));
#pragma unroll
            for (int k = 0; k < VZ; ++k) c[k] = ti.v[k];
        } else {
            const SVec<unsigned, VZ> ti = ld_stream(reinterpret_cast<const SVec<unsigned, VZ>*>((const unsigned*)cf.idx + p));
#pragma unroll
            for (int k = 0; k < VZ; ++k) c[k] = ti.v[k];
        }
        table_rows<A, VZ, ONE>(cf.a, c, a, b);
    }
}

// curl of the voltage update for the vector z0..z0+VZ-1 at (x, y), p = its
// element offset: reads 3 + 4 current components (x-1, y-1 neighbours; the
// z-1 neighbour of the first cell, zero at z = 0)
template <typename S, int VZ>
__device__ __forceinline__ void volt_curl(const S* __restrict__ I, const Dims& d, const Scales<typename StoreTraits<S>::A>& sc,
                                          int x, int y, int z0, long long p, typename StoreTraits<S>::A curl[3][VZ]) {
    typedef StoreTraits<S> T;
    typedef typename T::A A;
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
#pragma unroll
    for (int k = 0; k < VZ; ++k) {
        const A i1_zm = k == 0 ? i1_zm0 : i1[k - 1];
        const A i0_zm = k == 0 ? i0_zm0 : i0[k - 1];
        curl[0][k] = ((i2[k] - i2_ym[k]) - i1[k]) + i1_zm;
        curl[1][k] = ((i0[k] - i0_zm) - i2[k]) + i2_xm[k];
        curl[2][k] = ((i1[k] - i1_xm[k]) - i0[k]) + i0_ym[k];
    }
}

// the voltage update of one vector from its curl: V = V*vv + vi*curl, with
// the UPML form for cells of the (at most two) slabs the vector touches
template <typename S, int CM, bool PML, int VZ>
__device__ __forceinline__ void volt_apply(S* __restrict__ V, const CoeffPtrs<typename StoreTraits<S>::A>& cf, const Dims& d,
                                           const Scales<typename StoreTraits<S>::A>& sc, const PmlSet& ps,
                                           const PmlPtrs<typename StoreTraits<S>::A>& pp, int x, int y, int z0, long long p,
                                           const typename StoreTraits<S>::A curl[3][VZ]) {
    typedef typename StoreTraits<S>::A A;
    A vv[3][VZ], vi[3][VZ];
    vload_coeffs<A, CM, VZ, OneRow<S>::value>(cf, p, d.sc, x, y, z0, vv, vi);
    const int valid = min(VZ, d.nz - z0);
    SlabHits hit{{-1, -1}, {0u, 0u}, {0, 0}};
    SVec<unsigned short, VZ> prow[2] = {};
    if (PML) {
        hit = slab_hits<VZ>(ps, x, y, z0, ~0u);
#pragma unroll
        for (int j = 0; j < 2; ++j)
            if (hit.s[j] >= 0 && pp.vtab) {
                const PmlSlab& b = ps.slab[hit.s[j]];
                prow[j] = *reinterpret_cast<const SVec<unsigned short, VZ>*>(pp.vidx + (hit.b[j] - b.off + b.ioff));
            }
    }
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        A v[VZ], out[VZ];
        vload_stream<S, VZ>(V + n * d.sc + p, v, sc.v[n]);
#pragma unroll
        for (int k = 0; k < VZ; ++k) {
            out[k] = v[k] * vv[n][k];
            out[k] = out[k] + vi[n][k] * curl[n][k];
        }
        if (PML) {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if (hit.s[j] < 0) continue;
                const PmlSlab& b = ps.slab[hit.s[j]];
                upml_lanes<A, VZ, OneRow<S>::value>(pp.vflux, pp.vv, pp.vvfo, pp.vvfn, pp.vtab, prow[j].v, hit.b[j] + n * b.sc, n, hit.m[j], v,
                                  vv[n], vi[n], curl[n], out);
            }
        }
        vstore<S, VZ>(V + n * d.sc + p, out, sc.iv[n], valid);
    }
}

// one vector of VZ cells at plane x, flat (y,z) offset q0 of the padded plane
template <typename S, int CM, bool PML, int VZ>
__device__ __forceinline__ void volt_vec_body(S* __restrict__ V, const S* __restrict__ I, const CoeffPtrs<typename StoreTraits<S>::A>& cf,
                                              const Dims& d, const Scales<typename StoreTraits<S>::A>& sc, const PmlSet& ps,
                                              const PmlPtrs<typename StoreTraits<S>::A>& pp, int x, long long q0) {
    typedef typename StoreTraits<S>::A A;
    const int y = (int)(q0 / d.sy);
    const int z0 = (int)(q0 - (long long)y * d.sy);
    if (y >= d.ny || z0 >= d.nz) return;
    const long long p = x * d.sx + q0;
    A curl[3][VZ];
    volt_curl<S, VZ>(I, d, sc, x, y, z0, p, curl);
    volt_apply<S, CM, PML, VZ>(V, cf, d, sc, ps, pp, x, y, z0, p, curl);
}

// Vectors one launch of k_volt_vec / k_curr_vec covers: planes x0 + blockIdx.y,
// rows y0..y1-1 (flat (y, z) index from y0 over the padded rows), and only
// vectors whose cells z0..z0+VZ-1 lie in [zlo, zhi).  With UPML slabs forming
// the shell around an interior box, the interior vectors run here without
// PML code and k_volt_shell / k_curr_shell take every other vector.
// bpp > 0 selects the banded order (a 1D grid): the rows are cut into bands
// of tpb tiles, and the launch sweeps band by band, each band over all its
// planes, so the x-1 (x+1) plane of a band is still in L2 when the next
// plane reads it again even where a whole plane would not fit.  A band-plane
// takes bpp blocks.  xb > 0 (XCD chunks of xb tiles): workgroups go to the
// eight XCDs in turn (block k to XCD k % 8), and block k of a band-plane
// takes tile (c * 8 + k % 8) * xb + o with s = k / 8, c = s / xb, o = s % xb:
// XCD j sweeps chunks j, j + 8, j + 16, ... of xb consecutive tiles, the
// same chunks in every plane, so the y-1 (y+1) rows and the x-1 (x+1) plane
// its blocks read again were loaded by the same XCD; bpp is then a multiple
// of 8 * xb.  xb >= tpb / 8 gives each XCD one contiguous sub-band.
// bpp = 0: a 2D grid, tile = blockIdx.x, plane = blockIdx.y.
struct VecRange {
    int x0, y0, y1, zlo, zhi;
    int bpp, tpb, nx, xb;
};

// tile and plane of block b (linear, ascending order) of a banded launch;
// false for the padding blocks of a band-plane
__device__ __forceinline__ bool band_tile(const VecRange& r, unsigned b, long long& tile, int& px) {
    const unsigned per = (unsigned)r.bpp * (unsigned)r.nx;
    const unsigned band = b / per;
    const unsigned rem = b - band * per;
    px = (int)(rem / (unsigned)r.bpp);
    const unsigned k = rem - (unsigned)px * (unsigned)r.bpp;
    unsigned t = k;
    if (r.xb) {
        const unsigned s = k >> 3, c = s / (unsigned)r.xb;
        t = (c * 8u + (k & 7u)) * (unsigned)r.xb + (s - c * (unsigned)r.xb);
    }
    if (t >= (unsigned)r.tpb) return false;
    tile = (long long)band * r.tpb + t;
    return true;
}

template <typename S, int CM, bool PML, int VZ>
__device__ __forceinline__ void volt_vec_range(S* __restrict__ V, const S* __restrict__ I, const CoeffPtrs<typename StoreTraits<S>::A>& cf,
                                               const Dims& d, const Scales<typename StoreTraits<S>::A>& sc, const PmlSet& ps,
                                               const PmlPtrs<typename StoreTraits<S>::A>& pp, const VecRange& r) {
    long long tile = blockIdx.x;
    int px = (int)blockIdx.y;
    if (r.bpp && !band_tile(r, blockIdx.x, tile, px)) return;
    const long long q0 = (long long)r.y0 * d.sy + (tile * blockDim.x + threadIdx.x) * VZ;
    const int y = (int)(q0 / d.sy);
    const int z0 = (int)(q0 - (long long)y * d.sy);
    if (y >= r.y1 || z0 < r.zlo || z0 + VZ > r.zhi) return;
    volt_vec_body<S, CM, PML, VZ>(V, I, cf, d, sc, ps, pp, r.x0 + px, q0);
}

// Without UPML code the voltage update fits 5 waves per SIMD for every
// storage type (the bound keeps FP16/BF16 at 96 VGPRs without spills),
// except 2-byte storage with factored coefficients: forced to 96 VGPRs it
// serialises its loads (graded 257^3 port: 132 us against 70 us for the
// current update at 127 VGPRs), so it gets 4 waves.  The current update would
// spill under the same bound and the UPML forms need more registers, so they
// get none.
template <typename S, int CM, int VZ>
__global__ void __launch_bounds__(256, (CM == CF_FACT16 && sizeof(S) == 2) ? 4 : 5)
k_volt_vec(S* __restrict__ V, const S* __restrict__ I, CoeffPtrs<typename StoreTraits<S>::A> cf, Dims d,
           Scales<typename StoreTraits<S>::A> sc, PmlSet ps, PmlPtrs<typename StoreTraits<S>::A> pp, VecRange r) {
    volt_vec_range<S, CM, false, VZ>(V, I, cf, d, sc, ps, pp, r);
}

template <typename S, int CM, int VZ>
__global__ void __launch_bounds__(256)
k_volt_vec_upml(S* __restrict__ V, const S* __restrict__ I, CoeffPtrs<typename StoreTraits<S>::A> cf, Dims d,
                Scales<typename StoreTraits<S>::A> sc, PmlSet ps, PmlPtrs<typename StoreTraits<S>::A> pp, VecRange r) {
    volt_vec_range<S, CM, true, VZ>(V, I, cf, d, sc, ps, pp, r);
}

// The same with a 5-wave register bound, for FP32/FP64 storage with full
// coefficient arrays: there it costs at most 4 spilled VGPRs and gains 2-3%
// (graded 257^3 port, FP32); the index forms spill 12 and lose as much, and
// 2-byte storage would spill 83, so they keep the unbounded kernel.
template <typename S, int CM, int VZ>
__global__ void __launch_bounds__(256, 5)
k_volt_vec_upml_b5(S* __restrict__ V, const S* __restrict__ I, CoeffPtrs<typename StoreTraits<S>::A> cf, Dims d,
                   Scales<typename StoreTraits<S>::A> sc, PmlSet ps, PmlPtrs<typename StoreTraits<S>::A> pp, VecRange r) {
    volt_vec_range<S, CM, true, VZ>(V, I, cf, d, sc, ps, pp, r);
}

// The vectors outside the interior box of a UPML shell, numbered
// (x < x0 planes | x1 <= x < xe planes | interior planes: rows y < y0 or
// y >= y1 | interior rows: vectors before vlo or from vhi on), each with
// vectors along z fastest.
struct ShellGrid {
    int x0, x1, xe;            // interior planes [x0, x1); planes end at xe
    int y0, y1, ny;            // interior rows [y0, y1) of ny
    int vlo, vhi, vn;          // interior vectors [vlo, vhi) of vn per row
    long long nA0, nA1, nB, nC;
};

// Vector t of the shell, in 32-bit arithmetic: the engine splits only shells
// of fewer than 2^31 vectors.
__host__ __device__ inline bool shell_vector(const ShellGrid& g, unsigned t, int vz, long long sy, int& x, long long& q0) {
    unsigned y, v;
    const unsigned vn = (unsigned)g.vn;
    const unsigned row = (unsigned)g.ny * vn;
    if (t < (unsigned)g.nA0) {
        x = (int)(t / row); const unsigned r = t - (unsigned)x * row;
        y = r / vn; v = r - y * vn;
    } else if ((t -= (unsigned)g.nA0) < (unsigned)g.nA1) {
        x = (int)(t / row); const unsigned r = t - (unsigned)x * row;
        x += g.x1;
        y = r / vn; v = r - y * vn;
    } else if ((t -= (unsigned)g.nA1) < (unsigned)g.nB) {
        const unsigned nyb = (unsigned)(g.y0 + g.ny - g.y1);
        const unsigned per = nyb * vn;
        x = (int)(t / per); const unsigned r = t - (unsigned)x * per;
        x += g.x0;
        const unsigned yy = r / vn;
        v = r - yy * vn;
        y = yy < (unsigned)g.y0 ? yy : (unsigned)g.y1 + (yy - (unsigned)g.y0);
    } else if ((t -= (unsigned)g.nB) < (unsigned)g.nC) {
        const unsigned nvz = (unsigned)(g.vlo + g.vn - g.vhi);
        const unsigned per = (unsigned)(g.y1 - g.y0) * nvz;
        x = (int)(t / per); const unsigned r = t - (unsigned)x * per;
        x += g.x0;
        const unsigned yy = r / nvz;
        const unsigned j = r - yy * nvz;
        y = (unsigned)g.y0 + yy;
        v = j < (unsigned)g.vlo ? j : (unsigned)g.vhi + (j - (unsigned)g.vlo);
    } else {
        return false;
    }
    q0 = (long long)y * sy + (long long)v * vz;
    return true;
}

template <typename S, int CM, int VZ>
__global__ void __launch_bounds__(256)
k_volt_shell(S* __restrict__ V, const S* __restrict__ I, CoeffPtrs<typename StoreTraits<S>::A> cf, Dims d,
             Scales<typename StoreTraits<S>::A> sc, PmlSet ps, PmlPtrs<typename StoreTraits<S>::A> pp, ShellGrid g) {
    int x;
    long long q0;
    if (!shell_vector(g, blockIdx.x * blockDim.x + threadIdx.x, VZ, d.sy, x, q0)) return;
    volt_vec_body<S, CM, true, VZ>(V, I, cf, d, sc, ps, pp, x, q0);
}

// curl of the current update for the vector z0..z0+VZ-1 at (x, y): reads 3
// + 4 voltage components (x+1, y+1 neighbours; the z+1 neighbour of the
// last cell only where it is a valid line)
template <typename S, int VZ>
__device__ __forceinline__ void curr_curl(const S* __restrict__ V, const Dims& d, const Scales<typename StoreTraits<S>::A>& sc,
                                          int z0, long long p, typename StoreTraits<S>::A curl[3][VZ]) {
    typedef StoreTraits<S> T;
    typedef typename T::A A;
    const long long c1 = d.sc, c2 = 2 * d.sc;
    A v0[VZ], v1[VZ], v2[VZ], v2_yp[VZ], v0_yp[VZ], v2_xp[VZ], v1_xp[VZ];
    vload<S, VZ>(V + p, v0, sc.v[0]);
    vload<S, VZ>(V + c1 + p, v1, sc.v[1]);
    vload<S, VZ>(V + c2 + p, v2, sc.v[2]);
    vload<S, VZ>(V + c2 + p + d.sy, v2_yp, sc.v[2]);
    vload<S, VZ>(V + p + d.sy, v0_yp, sc.v[0]);
    vload<S, VZ>(V + c2 + p + d.sx, v2_xp, sc.v[2]);
    vload<S, VZ>(V + c1 + p + d.sx, v1_xp, sc.v[1]);
    const bool tail = z0 + VZ < d.nz;
    const A v1_zpl = tail ? T::load(V + c1 + p + VZ, sc.v[1]) : A(0);
    const A v0_zpl = tail ? T::load(V + p + VZ, sc.v[0]) : A(0);
#pragma unroll
    for (int k = 0; k < VZ; ++k) {
        const A v1_zp = k == VZ - 1 ? v1_zpl : v1[k + 1];
        const A v0_zp = k == VZ - 1 ? v0_zpl : v0[k + 1];
        curl[0][k] = ((v2[k] - v2_yp[k]) - v1[k]) + v1_zp;
        curl[1][k] = ((v0[k] - v0_zp) - v2[k]) + v2_xp[k];
        curl[2][k] = ((v1[k] - v1_xp[k]) - v0[k]) + v0_yp[k];
    }
}

// the current update of one vector from its curl, UPML form included
template <typename S, int CM, bool PML, int VZ>
__device__ __forceinline__ void curr_apply(S* __restrict__ I, const CoeffPtrs<typename StoreTraits<S>::A>& cf, const Dims& d,
                                           const Scales<typename StoreTraits<S>::A>& sc, const PmlSet& ps,
                                           const PmlPtrs<typename StoreTraits<S>::A>& pp, int x, int y, int z0, long long p,
                                           const typename StoreTraits<S>::A curl[3][VZ]) {
    typedef typename StoreTraits<S>::A A;
    A ii[3][VZ], iv[3][VZ];
    vload_coeffs<A, CM, VZ, OneRow<S>::value>(cf, p, d.sc, x, y, z0, ii, iv);
    const int valid = min(VZ, d.nz - 1 - z0);
    // cells at z >= nz-1 are not updated: keep them out of the slab masks too
    const unsigned upd = valid < VZ ? (1u << valid) - 1u : ~0u;
    SlabHits hit{{-1, -1}, {0u, 0u}, {0, 0}};
    SVec<unsigned short, VZ> prow[2] = {};
    if (PML) {
        hit = slab_hits<VZ>(ps, x, y, z0, upd);
#pragma unroll
        for (int j = 0; j < 2; ++j)
            if (hit.s[j] >= 0 && pp.ctab) {
                const PmlSlab& b = ps.slab[hit.s[j]];
                prow[j] = *reinterpret_cast<const SVec<unsigned short, VZ>*>(pp.cidx + (hit.b[j] - b.off + b.ioff));
            }
    }
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        A c[VZ], out[VZ];
        vload_stream<S, VZ>(I + n * d.sc + p, c, sc.c[n]);
#pragma unroll
        for (int k = 0; k < VZ; ++k) {
            out[k] = c[k] * ii[n][k];
            out[k] = out[k] + iv[n][k] * curl[n][k];
        }
        if (PML) {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if (hit.s[j] < 0) continue;
                const PmlSlab& b = ps.slab[hit.s[j]];
                upml_lanes<A, VZ, OneRow<S>::value>(pp.cflux, pp.ii, pp.iifo, pp.iifn, pp.ctab, prow[j].v, hit.b[j] + n * b.sc, n, hit.m[j], c,
                                  ii[n], iv[n], curl[n], out);
            }
        }
        vstore<S, VZ>(I + n * d.sc + p, out, sc.ic[n], valid);
    }
}

template <typename S, int CM, bool PML, int VZ>
__device__ __forceinline__ void curr_vec_body(S* __restrict__ I, const S* __restrict__ V, const CoeffPtrs<typename StoreTraits<S>::A>& cf,
                                              const Dims& d, const Scales<typename StoreTraits<S>::A>& sc, const PmlSet& ps,
                                              const PmlPtrs<typename StoreTraits<S>::A>& pp, int x, long long q0) {
    typedef typename StoreTraits<S>::A A;
    const int y = (int)(q0 / d.sy);
    const int z0 = (int)(q0 - (long long)y * d.sy);
    if (y >= d.ny - 1 || z0 >= d.nz - 1) return;
    const long long p = x * d.sx + q0;
    A curl[3][VZ];
    curr_curl<S, VZ>(V, d, sc, z0, p, curl);
    curr_apply<S, CM, PML, VZ>(I, cf, d, sc, ps, pp, x, y, z0, p, curl);
}

// REV: blocks walk planes and tiles in descending order, so the x+1 and
// y+1 neighbours a block reads were loaded by the blocks just before it
// (the mirror of the voltage update's x-1 / y-1 neighbours)
template <typename S, int CM, bool PML, int VZ, bool REV = false>
__device__ __forceinline__ void curr_vec_range(S* __restrict__ I, const S* __restrict__ V, const CoeffPtrs<typename StoreTraits<S>::A>& cf,
                                               const Dims& d, const Scales<typename StoreTraits<S>::A>& sc, const PmlSet& ps,
                                               const PmlPtrs<typename StoreTraits<S>::A>& pp, const VecRange& r) {
    const unsigned bx = REV ? gridDim.x - 1 - blockIdx.x : blockIdx.x;
    long long tile = bx;
    int px = REV ? (int)(gridDim.y - 1 - blockIdx.y) : (int)blockIdx.y;
    if (r.bpp && !band_tile(r, bx, tile, px)) return;
    const long long q0 = (long long)r.y0 * d.sy + (tile * blockDim.x + threadIdx.x) * VZ;
    const int y = (int)(q0 / d.sy);
    const int z0 = (int)(q0 - (long long)y * d.sy);
    if (y >= r.y1 || z0 < r.zlo || z0 + VZ > r.zhi) return;
    curr_vec_body<S, CM, PML, VZ>(I, V, cf, d, sc, ps, pp, r.x0 + px, q0);
}

// FP32 storage walks the planes in descending order (512^3: 0.946 -> 0.907 ms
// per update); the UPML form and 2-byte storage measured 2-3% slower that way
template <typename S, int CM, int VZ>
__global__ void __launch_bounds__(256)
k_curr_vec(S* __restrict__ I, const S* __restrict__ V, CoeffPtrs<typename StoreTraits<S>::A> cf, Dims d,
           Scales<typename StoreTraits<S>::A> sc, PmlSet ps, PmlPtrs<typename StoreTraits<S>::A> pp, VecRange r) {
    curr_vec_range<S, CM, false, VZ, sizeof(S) == 4>(I, V, cf, d, sc, ps, pp, r);
}

template <typename S, int CM, int VZ>
__global__ void __launch_bounds__(256)
k_curr_vec_upml(S* __restrict__ I, const S* __restrict__ V, CoeffPtrs<typename StoreTraits<S>::A> cf, Dims d,
                Scales<typename StoreTraits<S>::A> sc, PmlSet ps, PmlPtrs<typename StoreTraits<S>::A> pp, VecRange r) {
    curr_vec_range<S, CM, true, VZ>(I, V, cf, d, sc, ps, pp, r);
}

template <typename S, int CM, int VZ>
__global__ void __launch_bounds__(256, 5)
k_curr_vec_upml_b5(S* __restrict__ I, const S* __restrict__ V, CoeffPtrs<typename StoreTraits<S>::A> cf, Dims d,
                   Scales<typename StoreTraits<S>::A> sc, PmlSet ps, PmlPtrs<typename StoreTraits<S>::A> pp, VecRange r) {
    curr_vec_range<S, CM, true, VZ>(I, V, cf, d, sc, ps, pp, r);
}

template <typename S, int CM, int VZ>
__global__ void __launch_bounds__(256)
k_curr_shell(S* __restrict__ I, const S* __restrict__ V, CoeffPtrs<typename StoreTraits<S>::A> cf, Dims d,
             Scales<typename StoreTraits<S>::A> sc, PmlSet ps, PmlPtrs<typename StoreTraits<S>::A> pp, ShellGrid g) {
    int x;
    long long q0;
    if (!shell_vector(g, blockIdx.x * blockDim.x + threadIdx.x, VZ, d.sy, x, q0)) return;
    curr_vec_body<S, CM, true, VZ>(I, V, cf, d, sc, ps, pp, x, q0);
}

// ---------------------------------------------------------------------------
// UPML on the current cells the main update never touches (the planes
// x = Nx-1, y = Ny-1 and z = Nz-1): Engine_Ext_UPML's pre and post hooks run
// over every slab cell (engine_ext_upml.cpp:141-234) while
// Engine::UpdateCurrents skips these, so for them pre + post alone give
//   I' = (k*I - fo*F) + fn*F,  F unchanged.
// They stay zero in a run from zero fields; values set there by SetCurr or
// fdtd_set_field evolve as in the reference.  One thread per face cell:
// face X (x = Nx-1, all y, z), face Y (y = Ny-1, x < Nx-1), face Z (z = Nz-1,
// x < Nx-1, y < Ny-1).
// ---------------------------------------------------------------------------
template <typename S>
__global__ void __launch_bounds__(256)
k_curr_upml_faces(S* __restrict__ I, Dims d, Scales<typename StoreTraits<S>::A> sc, PmlSet ps, PmlPtrs<typename StoreTraits<S>::A> pp) {
    typedef StoreTraits<S> T;
    typedef typename T::A A;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long fx = (long long)d.ny * d.nz, fy = (long long)(d.nx - 1) * d.nz, fz = (long long)(d.nx - 1) * (d.ny - 1);
    int x, y, z;
    if (t < fx) {
        x = d.nx - 1; y = (int)(t / d.nz); z = (int)(t - (long long)y * d.nz);
    } else if ((t -= fx) < fy) {
        y = d.ny - 1; x = (int)(t / d.nz); z = (int)(t - (long long)x * d.nz);
    } else if ((t -= fy) < fz) {
        z = d.nz - 1; x = (int)(t / (d.ny - 1)); y = (int)(t - (long long)x * (d.ny - 1));
    } else {
        return;
    }
    for (int s = 0; s < ps.count; ++s) {
        const PmlSlab& b = ps.slab[s];
        const int lx = x - b.s[0], ly = y - b.s[1], lz = z - b.s[2];
        if ((unsigned)lx >= (unsigned)b.n[0] || (unsigned)ly >= (unsigned)b.n[1] || (unsigned)lz >= (unsigned)b.n[2]) continue;
        const long long l = ((long long)lx * b.n[1] + ly) * b.nzp + (z - b.za);   // component-0 slab offset from b.off
        const long long p = x * d.sx + (long long)y * d.sy + z;
        for (int n = 0; n < 3; ++n) {
            const long long q = b.off + n * b.sc + l;
            A k, fo, fn;
            if (pp.ctab) {
                const A* r = pp.ctab + (size_t)pp.cidx[l + b.ioff] * 12;
                k = r[n]; fo = r[4 + n]; fn = r[8 + n];
            } else {
                k = pp.ii[q]; fo = pp.iifo[q]; fn = pp.iifn[q];
            }
            const A F = pp.cflux[q];
            const A v = T::load(I + n * d.sc + p, sc.c[n]);
            const A h = k * v - fo * F;
            T::store(I + n * d.sc + p, h + fn * F, sc.ic[n]);
        }
        return;
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

// ---------------------------------------------------------------------------
// Mur ABC, first order (Engine_Ext_Mur_ABC, engine_ext_mur_abc.cpp:83-168).
// For the two tangential components t of the boundary plane `line` of axis
// ny, with the neighbouring plane `shift` one line inside:
//   pre   (before the voltage update)  buf = V_t(shift) - k_t * V_t(line)
//   post  (after the update)           buf = buf + k_t * V_t(shift)
//   apply                              V_t(line) = buf
// from timestep start_ts on (before it, nothing: excitations on the plane).
// ---------------------------------------------------------------------------
struct MurDev {
    int ny, nyP, nyPP;
    int line, shift;
    int na, nb;                // lines along nyP and nyPP
    unsigned start_ts;
    long long sny;             // element stride along ny
    const void* k;             // [2][na][nb] coefficients (nyP, nyPP)
    void* buf;                 // [2][na][nb]
};

template <typename S>
__global__ void k_mur(S* V, const unsigned* __restrict__ step, MurDev m, Dims d, Scales<typename StoreTraits<S>::A> sc, int phase) {
    typedef StoreTraits<S> T;
    typedef typename T::A A;
    const long long plane = (long long)m.na * m.nb;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= plane || *step < m.start_ts) return;
    const int a = (int)(i / m.nb), b = (int)(i - (long long)a * m.nb);
    int pos[3];
    pos[m.ny] = m.line;
    pos[m.nyP] = a;
    pos[m.nyPP] = b;
    const long long p = pos[0] * d.sx + pos[1] * d.sy + pos[2];
    const long long ps = p + (long long)(m.shift - m.line) * m.sny;
    const A* k = (const A*)m.k;
    A* buf = (A*)m.buf;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int c = t == 0 ? m.nyP : m.nyPP;
        const long long j = t * plane + i;
        S* F = V + c * d.sc;
        if (phase == 0) {
            const A kv = k[j] * T::load(F + p, sc.v[c]);
            buf[j] = T::load(F + ps, sc.v[c]) - kv;
        } else if (phase == 1) {
            const A kv = k[j] * T::load(F + ps, sc.v[c]);
            buf[j] = buf[j] + kv;
        } else {
            T::store(F + p, buf[j], sc.iv[c]);
        }
    }
}

// ---------------------------------------------------------------------------
// Lumped RLC (Engine_Ext_LumpedRLC, engine_ext_lumpedRLC.cpp:81-149).  Per
// cell: the parallel inductor current Il, the last three node voltages Vd and
// the last three series currents J.  The reference rotates the three-slot
// buffers once per step (Vd in the pre hook, J in the apply hook); here the
// slot of logical entry k at timestep n is (k - (n + 1)) mod 3.
//   pre:   Il += (i2v * ilv) * Vd[1]
//   apply: Vd[0] = V; Vd[0] = vvd * ((((Vd[0] - Il) + vv2*Vd[2]) + vj1*J[1]) + vj2*J[2]);
//          J[0] = (ib0*(Vd[0] - Vd[2]) - (b1*ib0)*J[1]) - (b2*ib0)*J[2];  V = Vd[0]
// ---------------------------------------------------------------------------
struct RlcDev {
    int count;
    const long long* off;      // [count] element offset incl. the component
    const int* comp;           // [count]
    const void* c;             // [9][count]: ilv, i2v, vv2, vj1, vj2, vvd, ib0, b1, b2
    void* il;                  // [count]
    void* vd;                  // [3][count]
    void* j;                   // [3][count]
};

__device__ __forceinline__ int rlc_slot(int k, unsigned n) {
    return (int)((unsigned)(k + 3) - (n + 1u) % 3u) % 3;
}

template <typename S>
__global__ void k_rlc(S* V, const unsigned* __restrict__ step, RlcDev r, Scales<typename StoreTraits<S>::A> sc, int apply) {
    typedef StoreTraits<S> T;
    typedef typename T::A A;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r.count) return;
    const unsigned n = *step;
    const A* c = (const A*)r.c;
    A* il = (A*)r.il;
    A* vd = (A*)r.vd;
    A* jj = (A*)r.j;
    const long long cn = r.count;
    const A* ilv = c; const A* i2v = c + cn; const A* vv2 = c + 2 * cn; const A* vj1 = c + 3 * cn;
    const A* vj2 = c + 4 * cn; const A* vvd = c + 5 * cn; const A* ib0 = c + 6 * cn; const A* b1 = c + 7 * cn; const A* b2 = c + 8 * cn;
    A* v0 = vd + rlc_slot(0, n) * cn;
    A* v1 = vd + rlc_slot(1, n) * cn;
    A* v2 = vd + rlc_slot(2, n) * cn;
    if (!apply) {
        il[i] = il[i] + (i2v[i] * ilv[i]) * v1[i];
        return;
    }
    A* j0 = jj + rlc_slot(0, n) * cn;
    A* j1 = jj + rlc_slot(1, n) * cn;
    A* j2 = jj + rlc_slot(2, n) * cn;
    const int comp = r.comp[i];
    S* F = V + r.off[i];
    A x = T::load(F, sc.v[comp]);
    A t = x - il[i];
    t = t + vv2[i] * v2[i];
    t = t + vj1[i] * j1[i];
    t = t + vj2[i] * j2[i];
    x = vvd[i] * t;
    v0[i] = x;
    A u = ib0[i] * (x - v2[i]);
    u = u - (b1[i] * ib0[i]) * j1[i];
    u = u - (b2[i] * ib0[i]) * j2[i];
    j0[i] = u;
    T::store(F, x, sc.iv[comp]);
}

// ---------------------------------------------------------------------------
// Steady-state detection, Engine_Ext_SteadyState::Apply2Voltages
// (FDTD/extensions/engine_ext_steadystate.cpp:51-117).  The hook runs first
// among the voltage apply hooks (priority +2e6): after the post-update hooks,
// before lumped RLC, Lorentz, Mur and the excitation.
//
// Every step records the voltage at each probe edge into a ring of two
// periods (slot = numTS mod 2p, FP64).  At numTS = k*p, k >= 2, the energy of
// the whole domain (CalcFastEnergy) and, per probe, the power of the "new"
// window and its squared difference to the "old" one give
//   diff = max(|E - E_last| / E_last, max over probes with power above 1% of
//              the largest: diff_pow / curr_pow), 1 when nothing is valid or
//              the value exceeds 1.
// The windows follow the reference's choice: rel = numTS mod 2p is 0 or p at
// these steps, both satisfy rel <= p, so new = slots [0, p), old = [p, 2p).
// ---------------------------------------------------------------------------
#define SS_MAX_PROBES 256
struct SsDev {
    int n;                     // probes, <= SS_MAX_PROBES
    unsigned p;                // period in timesteps
    const long long* off;      // [n] element offset incl. the component
    const int* comp;           // [n]
    double* rec;               // [n][2p]
    double* state;             // [0] last diff, [1] last energy
};

template <typename S>
__global__ void k_ss_record(const S* __restrict__ V, const unsigned* __restrict__ step, SsDev d, Scales<typename StoreTraits<S>::A> sc) {
    typedef StoreTraits<S> T;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.n) return;
    const unsigned two_p = 2u * d.p;
    const unsigned rel = *step % two_p;
    d.rec[(size_t)i * two_p + rel] = (double)T::load(V + d.off[i], sc.v[d.comp[i]]);
}

// one block of 256: reduces the energy partials of k_energy_partial, then the
// per-probe window sums (each a serial loop in the reference's order, so the
// probe terms are exact for identical records), then the decision in thread 0
static __global__ void __launch_bounds__(256) k_ss_eval(const double* __restrict__ part, int nblk, SsDev d, double eps0, double mue0) {
    __shared__ double se[256], sh[256];
    __shared__ double cpow[SS_MAX_PROBES], dpow[SS_MAX_PROBES];
    double e = 0.0, h = 0.0;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) { e += part[2 * i]; h += part[2 * i + 1]; }
    se[threadIdx.x] = e;
    sh[threadIdx.x] = h;
    const unsigned p = d.p;
    const unsigned new_pos = 0, old_pos = p;
    for (int i = threadIdx.x; i < d.n; i += blockDim.x) {
        const double* b = d.rec + (size_t)i * 2u * p;
        double c = 0.0, q = 0.0;
        for (unsigned t = 0; t < p; ++t) {
            const double bn = b[t + new_pos], bo = b[t + old_pos];
            c += bn * bn;
            q += (bo - bn) * (bo - bn);
        }
        cpow[i] = c;
        dpow[i] = q;
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) { se[threadIdx.x] += se[threadIdx.x + s]; sh[threadIdx.x] += sh[threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double energy = eps0 * se[0] + mue0 * sh[0];
    const double last = d.state[1];
    bool no_valid = true;
    double diff = 0.0;
    if (last > 0) {
        diff = fabs(energy - last) / last;
        no_valid = false;
    }
    double max_pow = 0.0;
    for (int i = 0; i < d.n; ++i) max_pow = max_pow < cpow[i] ? cpow[i] : max_pow;
    for (int i = 0; i < d.n; ++i)
        if (cpow[i] > max_pow * 1e-2) {
            const double r = dpow[i] / cpow[i];
            diff = diff < r ? r : diff;
            no_valid = false;
        }
    if (no_valid || diff > 1) diff = 1;
    d.state[0] = diff;
    d.state[1] = energy;
}

struct AdeDev {
    int count;
    const long long* off;          // [count] element offset of position (component 0)
    const void* v_int; const void* v_ext; const void* v_lor;   // [3][count]
    const void* i_int; const void* i_ext; const void* i_lor;
    void* volt_ade; void* volt_lor; void* curr_ade; void* curr_lor;  // state [3][count]
    int volt_on, curr_on, volt_lor_on, curr_lor_on;
    // coefficient tables (volt, curr; null: the lists above): per entry a
    // uint16 row of distinct {int0..2, _, ext0..2, _, lor0..2, _}
    const unsigned short* cidx[2];
    const void* ctab[2];
    // fused hooks: per entry, bit n set when component n shares its location
    // with a TFSF or excitation entry of the half step (volt, curr); such
    // components are applied through the location's hook program
    const unsigned char* shared[2];
};

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

#define MAX_TFSF_ROUNDS 4
struct TfsfDev {
    int count;
    int length, period, max_delay;
    // entries are grouped into rounds with distinct field locations; rounds
    // run one after another so repeated locations add in list order
    int rounds;
    int rstart[MAX_TFSF_ROUNDS + 1];
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
__global__ void k_tfsf(S* F, const unsigned* __restrict__ step, TfsfDev t, Scales<typename StoreTraits<S>::A> sc, int is_curr,
                       int first, int end) {
    typedef StoreTraits<S> T;
    typedef typename T::A A;
    const int i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= end) return;
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
// Fused sparse hooks of one half step: the TFSF injection, the dispersive
// apply (V -= J) and the soft excitation of every location they touch, in
// one launch instead of one per hook (and per TFSF round).  The host groups
// the entries by field location (HookProg: one thread per location, its ops
// in the reference's hook order -- TFSF in list order, then the ADE orders,
// then the excitation in list order), so a location's sum order is the
// reference's; the value is rounded to the storage type after every op, as
// the separate hooks store it.  The same launch runs the ADE pre-update of
// the next half step (Engine_Ext_LorentzMaterial::DoPre*Updates) for the
// other field, which no hook of this half step changes: pre_next = 1.
// ---------------------------------------------------------------------------
#define MAX_ADE_ORDERS 4
struct AdeSet {
    int n;
    AdeDev o[MAX_ADE_ORDERS];
};

struct HookProg {
    int ntarget;
    const long long* off;          // [ntarget] element offset incl. component
    const int* start;              // [ntarget + 1] op ranges
    const unsigned* ops;           // type << 30 | index: 0 TFSF entry, 1 ADE (order << 26 | comp*count + entry), 2 excitation entry
};

template <typename S>
__device__ __forceinline__ typename StoreTraits<S>::A storage_round(typename StoreTraits<S>::A v, typename StoreTraits<S>::A s,
                                                                    typename StoreTraits<S>::A is) {
    if constexpr (sizeof(S) >= 4) {
        return v;
    } else {
        S t;
        StoreTraits<S>::store(&t, v, is);
        return StoreTraits<S>::load(&t, s);
    }
}

// Engine_Ext_LorentzMaterial::DoPre{Voltage,Current}UpdatesImpl for
// component n of entry i: J (and J_L) from the field value f and the
// current J value j
template <typename S>
__device__ __forceinline__ void ade_pre_comp(const AdeDev& a, int i, int n, typename StoreTraits<S>::A f, typename StoreTraits<S>::A j,
                                             int is_curr) {
    typedef typename StoreTraits<S>::A A;
    const A* kint = (const A*)(is_curr ? a.i_int : a.v_int);
    const A* kext = (const A*)(is_curr ? a.i_ext : a.v_ext);
    const A* klor = (const A*)(is_curr ? a.i_lor : a.v_lor);
    A* J = (A*)(is_curr ? a.curr_ade : a.volt_ade);
    A* JL = (A*)(is_curr ? a.curr_lor : a.volt_lor);
    const int lor_on = is_curr ? a.curr_lor_on : a.volt_lor_on;
    const A* row = a.ctab[is_curr] ? (const A*)a.ctab[is_curr] + 12 * (size_t)a.cidx[is_curr][i] : nullptr;
    const long long q = (long long)n * a.count + i;
    const A ki = row ? row[n] : kint[q];
    const A ke = row ? row[4 + n] : kext[q];
    if (lor_on) {
        A jl = JL[q];
        jl = jl + (row ? row[8 + n] : klor[q]) * j;
        j = j * ki;
        j = j + ke * (f - jl);
        JL[q] = jl;
        J[q] = j;
    } else {
        j = j * ki;
        j = j + ke * f;
        J[q] = j;
    }
}

// Apply (V -= J) of entry i's unshared components and, with pre, their
// pre-update of the next half step (the value just stored, rounded to the
// storage type, and J are its inputs).  Every load of the entry comes
// before its first store, so the components' memory round trips overlap
// (J and the field are separate arrays the compiler cannot tell apart).
template <typename S>
__device__ __forceinline__ void ade_apply_entry(S* __restrict__ F, const AdeDev& a, int i, long long sc_field,
                                                const Scales<typename StoreTraits<S>::A>& sc, int is_curr, bool pre) {
    typedef StoreTraits<S> T;
    typedef typename T::A A;
    const unsigned m = a.shared[is_curr] ? a.shared[is_curr][i] : 0u;
    if (m == 7u) return;
    A* J = (A*)(is_curr ? a.curr_ade : a.volt_ade);
    A* JL = (A*)(is_curr ? a.curr_lor : a.volt_lor);
    const int lor_on = pre && (is_curr ? a.curr_lor_on : a.volt_lor_on);
    const long long off = a.off[i];
    const A* row = pre && a.ctab[is_curr] ? (const A*)a.ctab[is_curr] + 12 * (size_t)a.cidx[is_curr][i] : nullptr;
    A f[3], j[3], jl[3], ki[3], ke[3], kl[3];
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        if (m & (1u << n)) continue;
        const long long q = (long long)n * a.count + i;
        f[n] = T::load(F + n * sc_field + off, is_curr ? sc.c[n] : sc.v[n]);
        j[n] = J[q];
        if (lor_on) jl[n] = JL[q];
        if (pre) {
            ki[n] = row ? row[n] : ((const A*)(is_curr ? a.i_int : a.v_int))[q];
            ke[n] = row ? row[4 + n] : ((const A*)(is_curr ? a.i_ext : a.v_ext))[q];
            if (lor_on) kl[n] = row ? row[8 + n] : ((const A*)(is_curr ? a.i_lor : a.v_lor))[q];
        }
    }
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        if (m & (1u << n)) continue;
        const long long q = (long long)n * a.count + i;
        const A s = is_curr ? sc.c[n] : sc.v[n];
        const A is = is_curr ? sc.ic[n] : sc.iv[n];
        A v = f[n] - j[n];
        T::store(F + n * sc_field + off, v, is);
        if (!pre) continue;
        v = storage_round<S>(v, s, is);
        A jj = j[n];
        if (lor_on) {
            A l = jl[n];
            l = l + kl[n] * jj;
            jj = jj * ki[n];
            jj = jj + ke[n] * (v - l);
            JL[q] = l;
            J[q] = jj;
        } else {
            jj = jj * ki[n];
            jj = jj + ke[n] * v;
            J[q] = jj;
        }
    }
}

// the pre-update of entry i from the field F: every component (mode 1) or
// only those that share their location with another hook (mode 2: the
// others were updated right after their apply, see k_hooks)
template <typename S>
__device__ __forceinline__ void ade_pre_entry(const S* __restrict__ F, const AdeDev& a, int i, long long sc_field,
                                              const Scales<typename StoreTraits<S>::A>& sc, int is_curr, int mode = 1) {
    typedef StoreTraits<S> T;
    typedef typename T::A A;
    const unsigned m = mode == 2 ? (a.shared[is_curr] ? a.shared[is_curr][i] : 0u) : 7u;
    if (!m) return;
    const A* J = (const A*)(is_curr ? a.curr_ade : a.volt_ade);
    const long long off = a.off[i];
    A f[3], j[3];
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        if (!(m & (1u << n))) continue;
        f[n] = T::load(F + n * sc_field + off, is_curr ? sc.c[n] : sc.v[n]);
        j[n] = J[(long long)n * a.count + i];
    }
#pragma unroll
    for (int n = 0; n < 3; ++n)
        if (m & (1u << n)) ade_pre_comp<S>(a, i, n, f[n], j[n], is_curr);
}

template <typename S>
__global__ void __launch_bounds__(256)
k_hooks(S* __restrict__ F, const S* __restrict__ G, const unsigned* __restrict__ step, HookProg hp, TfsfDev t, ExcDev e, AdeSet ade,
        Scales<typename StoreTraits<S>::A> sc, long long sc_field, int is_curr, int pre_next, int pre_self) {
    typedef StoreTraits<S> T;
    typedef typename T::A A;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < hp.ntarget) {
        const unsigned numTS = *step;
        const long long off = hp.off[i];
        const int n = (int)(off / sc_field);
        const A s = is_curr ? sc.c[n] : sc.v[n];
        const A is = is_curr ? sc.ic[n] : sc.iv[n];
        A v = T::load(F + off, s);
        for (int k = hp.start[i]; k < hp.start[i + 1]; ++k) {
            const unsigned op = hp.ops[k];
            const unsigned idx = op & 0x3fffffffu;
            switch (op >> 30) {
            case 0: {   // TFSF (k_tfsf)
                const A* sig = (const A*)t.sig;
                const A* amp = (const A*)t.amp;
                const A* delta = (const A*)t.delta;
                auto lookup = [&](unsigned m) -> unsigned {
                    unsigned L;
                    if (numTS < m) L = 0;
                    else if ((numTS - m >= (unsigned)t.length) && (t.period == 0)) L = 0;
                    else L = numTS - m;
                    if (t.period > 0) L = L % (unsigned)t.period;
                    return L;
                };
                const unsigned dl = (unsigned)t.delay[idx];
                const A second = delta[idx] * amp[idx] * sig[lookup(dl + 1)];
                const double r = (double)v + (1.0 - (double)delta[idx]) * (double)amp[idx] * (double)sig[lookup(dl)] + (double)second;
                v = (A)r;
                break;
            }
            case 1: {   // dispersive apply of a shared location (k_ade_apply)
                const AdeDev& a = ade.o[idx >> 26];
                const A* J = (const A*)(is_curr ? a.curr_ade : a.volt_ade);
                v = v - J[idx & 0x3ffffffu];
                break;
            }
            default: {  // soft excitation (k_excite)
                int p = (int)numTS + 1;
                if (e.period > 0) p = e.period;
                int kk = (int)numTS - e.delay[idx];
                kk *= (kk > 0);
                kk %= p;
                kk *= (kk < e.length);
                v = v + ((const A*)e.amp)[idx] * ((const A*)e.sig)[kk];
                break;
            }
            }
            v = storage_round<S>(v, s, is);
        }
        T::store(F + off, v, is);
    }
    // dispersive apply of the locations no other hook touches, per entry;
    // with pre_self, the same thread then runs the pre-update of this field's
    // next half step for them: no other hook of this or the other field's
    // half step touches these components again before it, so the value just
    // stored (rounded to the storage type) and J are its inputs
    for (int o = 0; o < ade.n; ++o) {
        const AdeDev& a = ade.o[o];
        if (i < a.count && (is_curr ? a.curr_on : a.volt_on)) ade_apply_entry<S>(F, a, i, sc_field, sc, is_curr, pre_self != 0);
    }
    if (pre_next) {
        for (int o = 0; o < ade.n; ++o) {
            const AdeDev& a = ade.o[o];
            if (i < a.count && (is_curr ? a.volt_on : a.curr_on)) ade_pre_entry<S>(G, a, i, sc_field, sc, !is_curr, pre_next);
        }
    }
}

// Engine_Ext_LorentzMaterial::DoPre{Voltage,Current}UpdatesImpl, engine_ext_lorentzmaterial.cpp:79-168
template <typename S>
__global__ void k_ade_pre(const S* __restrict__ F, AdeDev a, long long sc_field, Scales<typename StoreTraits<S>::A> sc, int is_curr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.count) return;
    ade_pre_entry<S>(F, a, i, sc_field, sc, is_curr);
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
__device__ __forceinline__ void step_end_body(const S* __restrict__ V, const S* __restrict__ I, unsigned* step, int advance,
                                              const ProbeDev* __restrict__ probes, int nprobes, const long long* __restrict__ e_off,
                                              const int* __restrict__ e_comp, const signed char* __restrict__ e_sign,
                                              const Scales<typename StoreTraits<S>::A>& sc, double dT) {
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

template <typename S>
__global__ void k_step_end(const S* __restrict__ V, const S* __restrict__ I, unsigned* step, int advance,
                           const ProbeDev* __restrict__ probes, int nprobes, const long long* __restrict__ e_off,
                           const int* __restrict__ e_comp, const signed char* __restrict__ e_sign,
                           Scales<typename StoreTraits<S>::A> sc, double dT) {
    step_end_body<S>(V, I, step, advance, probes, nprobes, e_off, e_comp, e_sign, sc, dT);
}

// ---------------------------------------------------------------------------
// Frequency-domain field accumulation over a box of mesh nodes
// (ProcessFieldsFD::Process, Common/processfields_fd.cpp:72-107) with node
// interpolation (Engine_Interface_FDTD::GetRawInterpolatedField and
// GetRawInterpolatedDualField, NODE_INTERPOLATE branches,
// FDTD/engine_interface_fdtd.cpp:72-97 and 134-158).  E = V / primary edge
// length, H = I / dual edge length (:257-266, :115-124).  Values and
// accumulators are FP64; the accumulator layout is [n_freq][3][nodes] of
// (re, im), nodes ordered z fastest.
// ---------------------------------------------------------------------------
#define MAX_DFT_FREQ 64

// Field values of dump boxes (Engine_Interface_FDTD,
// FDTD/engine_interface_fdtd.cpp:60-198,257-288), for the dump types of
// ProcessFields (Common/processfields.h:43): 0 E, 1 H, 2 J, 3 rot H, 4 D,
// 5 B, and the interpolation types of Engine_Interface_Base: 0 none, 1 node,
// 2 cell.  Values are FP64; the material arrays hold the operator's averaged
// absolute epsilon, kappa and mu per edge as float, [3][Nx][Ny][Nzp] in the
// padded field layout (null when not set: those types then give 0, as the
// reference does without stored materials).
struct DftAxes {
    const double* len;     // [Nx+Ny+Nz] primary edge lengths in metres
    const double* dlen;    // [Nx+Ny+Nz] dual edge lengths
    int off[3];
    int N[3];
};

struct DumpMat {
    const float* eps;
    const float* kappa;
    const float* mu;
};

struct DftDev {
    int type;              // dump type 0..5
    int interp;            // 0 none, 1 node, 2 cell
    int s[3], n[3];        // first node and node count per axis
    long long nodes;
    int n_freq;
    const double* freqs;
    double2* acc;          // [n_freq][3][nodes]
};

__device__ __forceinline__ bool dump_is_dual(int type) { return type == 1 || type == 5; }

template <typename S>
__device__ __forceinline__ double fval(const S* __restrict__ F, const Dims& d, int n, int x, int y, int z, typename StoreTraits<S>::A scale) {
    return (double)StoreTraits<S>::load(F + n * d.sc + x * d.sx + y * d.sy + z, scale);
}

// GetRawField (type: 0 E, 1 J, 2 rot H, 3 D)
template <typename S>
__device__ __forceinline__ double raw_primary(const S* __restrict__ V, const S* __restrict__ I, const Dims& d,
                                              const Scales<typename StoreTraits<S>::A>& sc, const DftAxes& ax, const DumpMat& m,
                                              int rt, int n, const int q[3]) {
    const long long c = (long long)n * d.sc + q[0] * d.sx + q[1] * d.sy + q[2];
    if (rt == 2) {
        const int nP = (n + 1) % 3, nPP = (n + 2) % 3;
        const double area = ax.dlen[ax.off[nP] + q[nP]] * ax.dlen[ax.off[nPP] + q[nPP]];
        double value = fval<S>(I, d, nPP, q[0], q[1], q[2], sc.c[nPP]);
        value -= fval<S>(I, d, nP, q[0], q[1], q[2], sc.c[nP]);
        int l[3] = {q[0], q[1], q[2]};
        if (q[nPP] > 0) {
            --l[nPP];
            value += fval<S>(I, d, nP, l[0], l[1], l[2], sc.c[nP]);
            ++l[nPP];
        }
        if (q[nP] > 0) {
            --l[nP];
            value -= fval<S>(I, d, nPP, l[0], l[1], l[2], sc.c[nPP]);
        }
        return value / area;
    }
    const double value = fval<S>(V, d, n, q[0], q[1], q[2], sc.v[n]);
    const double delta = ax.len[ax.off[n] + q[n]];
    if (rt == 0 && delta != 0.0) return value / delta;
    if (rt == 1 && m.kappa && delta != 0.0) return value * (double)m.kappa[c] / delta;
    if (rt == 3 && m.eps && delta != 0.0) return value * (double)m.eps[c] / delta;
    return 0.0;
}

// GetRawDualField (type: 0 H, 1 B)
template <typename S>
__device__ __forceinline__ double raw_dual(const S* __restrict__ I, const Dims& d, const Scales<typename StoreTraits<S>::A>& sc,
                                           const DftAxes& ax, const DumpMat& m, int rt, int n, const int q[3]) {
    const long long c = (long long)n * d.sc + q[0] * d.sx + q[1] * d.sy + q[2];
    const double value = fval<S>(I, d, n, q[0], q[1], q[2], sc.c[n]);
    const double delta = ax.dlen[ax.off[n] + q[n]];
    if (rt == 0 && delta != 0.0) return value / delta;
    if (rt == 1 && m.mu && delta != 0.0) return value * (double)m.mu[c] / delta;
    return 0.0;
}

// the dump value of `type` with interpolation `interp` at node pos
// (GetRawInterpolatedField / GetRawInterpolatedDualField)
template <typename S>
__device__ __forceinline__ void dump_value(const S* __restrict__ V, const S* __restrict__ I, const Dims& d,
                                           const Scales<typename StoreTraits<S>::A>& sc, const DftAxes& ax, const DumpMat& m,
                                           int type, int interp, const int pos[3], double out[3]) {
    if (!dump_is_dual(type)) {
        const int rt = type == 0 ? 0 : (type == 2 ? 1 : (type == 3 ? 2 : 3));
        if (interp == 2) {
            const bool edge = pos[0] == ax.N[0] - 1 || pos[1] == ax.N[1] - 1 || pos[2] == ax.N[2] - 1;
            for (int n = 0; n < 3; ++n) {
                const int nP = (n + 1) % 3, nPP = (n + 2) % 3;
                if (edge) { out[n] = 0.0; continue; }
                int q[3] = {pos[0], pos[1], pos[2]};
                double o = raw_primary<S>(V, I, d, sc, ax, m, rt, n, q);
                ++q[nP];
                o += raw_primary<S>(V, I, d, sc, ax, m, rt, n, q);
                ++q[nPP];
                o += raw_primary<S>(V, I, d, sc, ax, m, rt, n, q);
                --q[nP];
                o += raw_primary<S>(V, I, d, sc, ax, m, rt, n, q);
                out[n] = o / 4.0;
            }
            return;
        }
        for (int n = 0; n < 3; ++n) {
            int q[3] = {pos[0], pos[1], pos[2]};
            if (interp == 0) { out[n] = raw_primary<S>(V, I, d, sc, ax, m, rt, n, q); continue; }
            if (ax.N[n] < 2) { out[n] = 0.0; continue; }
            if (pos[n] == ax.N[n] - 1) {           // upper bound: the lower value only
                --q[n];
                out[n] = raw_primary<S>(V, I, d, sc, ax, m, rt, n, q);
                continue;
            }
            const double delta = ax.len[ax.off[n] + pos[n]];
            out[n] = raw_primary<S>(V, I, d, sc, ax, m, rt, n, q);
            if (delta == 0.0) { out[n] = 0.0; continue; }
            if (pos[n] == 0) continue;             // lower bound: the upper value only
            --q[n];
            const double deltaDown = ax.len[ax.off[n] + q[n]];
            const double rel = delta / (delta + deltaDown);
            out[n] = out[n] * (1.0 - rel) + raw_primary<S>(V, I, d, sc, ax, m, rt, n, q) * rel;
        }
        return;
    }
    const int rt = type == 1 ? 0 : 1;
    if (interp == 1) {
        const bool edge = pos[0] == ax.N[0] - 1 || pos[1] == ax.N[1] - 1 || pos[2] == ax.N[2] - 1;
        for (int n = 0; n < 3; ++n) {
            const int nP = (n + 1) % 3, nPP = (n + 2) % 3;
            if (edge || pos[nP] == 0 || pos[nPP] == 0) { out[n] = 0.0; continue; }
            int q[3] = {pos[0], pos[1], pos[2]};
            double o = raw_dual<S>(I, d, sc, ax, m, rt, n, q);
            --q[nP];
            o += raw_dual<S>(I, d, sc, ax, m, rt, n, q);
            --q[nPP];
            o += raw_dual<S>(I, d, sc, ax, m, rt, n, q);
            ++q[nP];
            o += raw_dual<S>(I, d, sc, ax, m, rt, n, q);
            out[n] = o / 4.0;
        }
        return;
    }
    for (int n = 0; n < 3; ++n) {
        int q[3] = {pos[0], pos[1], pos[2]};
        out[n] = raw_dual<S>(I, d, sc, ax, m, rt, n, q);
        if (interp == 0) continue;
        if (pos[n] >= ax.N[n] - 1) { out[n] = 0.0; continue; }     // outer boundary
        const double delta = ax.dlen[ax.off[n] + pos[n]];
        ++q[n];
        const double deltaUp = ax.dlen[ax.off[n] + q[n]];
        const double rel = delta / (delta + deltaUp);
        out[n] = out[n] * (1.0 - rel) + raw_dual<S>(I, d, sc, ax, m, rt, n, q) * rel;
    }
}

template <typename S>
__global__ void __launch_bounds__(256)
k_dft(const S* __restrict__ V, const S* __restrict__ I, Dims d, Scales<typename StoreTraits<S>::A> sc, DftAxes ax, DumpMat m, DftDev b,
      double t, double g) {
    __shared__ double2 ph[MAX_DFT_FREQ];
    for (int k = threadIdx.x; k < b.n_freq; k += blockDim.x) {
        double sn, cs;
        sincos(-2.0 * M_PI * b.freqs[k] * t, &sn, &cs);
        // exp(-2 pi i f T) * 2 * (dT * FD_Interval), in the reference's order
        ph[k] = make_double2((cs * 2.0) * g, (sn * 2.0) * g);
    }
    __syncthreads();
    const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= b.nodes) return;
    const int k = (int)(node % b.n[2]);
    const long long r = node / b.n[2];
    const int j = (int)(r % b.n[1]);
    const int i = (int)(r / b.n[1]);
    const int pos[3] = {b.s[0] + i, b.s[1] + j, b.s[2] + k};
    double out[3];
    dump_value<S>(V, I, d, sc, ax, m, b.type, b.interp, pos, out);
    for (int f = 0; f < b.n_freq; ++f) {
        const double2 p = ph[f];
        for (int n = 0; n < 3; ++n) {
            double2* a = b.acc + ((long long)f * 3 + n) * b.nodes + node;
            double2 v = *a;
            v.x += out[n] * p.x;
            v.y += out[n] * p.y;
            *a = v;
        }
    }
}

// Time-domain field dump of a node box (ProcessFieldsTD::Process via
// ProcessFields::CalcField, Common/processfields.cpp:280-400): the dump
// value rounded to float, [3][nodes] with nodes z fastest.
template <typename S>
__global__ void __launch_bounds__(256)
k_snapshot(const S* __restrict__ V, const S* __restrict__ I, Dims d, Scales<typename StoreTraits<S>::A> sc, DftAxes ax, DumpMat m, DftDev b,
           float* __restrict__ out) {
    const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= b.nodes) return;
    const int k = (int)(node % b.n[2]);
    const long long r = node / b.n[2];
    const int j = (int)(r % b.n[1]);
    const int i = (int)(r / b.n[1]);
    const int pos[3] = {b.s[0] + i, b.s[1] + j, b.s[2] + k};
    double v[3];
    dump_value<S>(V, I, d, sc, ax, m, b.type, b.interp, pos, v);
    for (int n = 0; n < 3; ++n) out[n * b.nodes + node] = (float)v[n];
}

// ---------------------------------------------------------------------------
// per-component max |value| of one field (physical units), over the padded
// component arrays (padding holds zeros): grid (blocks, 3), one partial per
// block and component.  Feeds the storage-scale pilot of FP16/BF16 handles
// without moving the fields to the host.
// ---------------------------------------------------------------------------
template <typename S>
__global__ void __launch_bounds__(256)
k_absmax(const S* __restrict__ F, long long sc, Scales<typename StoreTraits<S>::A> s, int is_curr, double* __restrict__ part) {
    typedef typename StoreTraits<S>::A A;
    constexpr int VZ = VecWidth<S>::value;
    __shared__ A sm[256];
    const int n = blockIdx.y;
    const S* p = F + n * sc;
    const A scale = is_curr ? s.c[n] : s.v[n];
    A m = A(0);
    for (long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * VZ; i < sc; i += (long long)gridDim.x * blockDim.x * VZ) {
        A v[VZ];
        vload<S, VZ>(p + i, v, scale);
#pragma unroll
        for (int k = 0; k < VZ; ++k) m = fmax(m, fabs(v[k]));
    }
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)n * gridDim.x + blockIdx.x] = (double)sm[0];
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

