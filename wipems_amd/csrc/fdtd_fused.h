// fdtd_fused.h -- one launch per timestep: voltage and current updates of a
// tile fused in an x-march, with ping-pong field buffers.
//
// A block owns a tile of TY rows (y) x TZ = 64*VZ cells (z) and a segment of
// x planes.  Marching x, it loads I_old(x) for its rows plus the row above
// into LDS, computes V_new(x) for its rows plus the row below (a halo row,
// recomputed here instead of read from the neighbour tile), then I_new(x-1)
// from V_new(x-1) (LDS, kept from the previous plane) and V_new(x)
// (registers).  The last plane of the segment is a halo plane: its V_new is
// computed but not stored.  Each owned cell therefore moves V and I in and
// out once per step (48 B in FP32 plus the coefficient index) instead of
// twice, which two update launches need.
//
// Results are bit-identical to the two-launch path: the same expressions in
// the same order, values rounded to the storage type exactly where the
// two-launch path stores them, soft excitation added after the voltage
// (current) update in list order.  Reads go to the "in" buffers and writes to
// the "out" buffers only, so recomputed halos see old values.  Cells the
// reference does not update (currents at x = Nx-1, y = Ny-1, z = Nz-1) are
// copied.  Used for domains without UPML, Lorentz or TFSF extensions.
#pragma once
#include "fdtd_device.h"

#define FUSED_MAX_EXC 64

// a value rounded to the storage type as a store and a load would round it
template <typename S>
__device__ __forceinline__ typename StoreTraits<S>::A round_store(typename StoreTraits<S>::A v, typename StoreTraits<S>::A s,
                                                                  typename StoreTraits<S>::A inv_s) {
    S t;
    StoreTraits<S>::store(&t, v, inv_s);
    return StoreTraits<S>::load(&t, s);
}

struct FusedGrid {
    int tiles_y, tiles_z, segs_x;
    int seg;                       // planes per x segment
    int nblocks, per_xcd;          // XCD-contiguous block order
};

// excitation signal index of Engine_Ext_Excitation::Apply2Voltages (as k_excite)
__device__ __forceinline__ int exc_index(int numTS, int delay, int period, int length) {
    int p = numTS + 1;
    if (period > 0) p = period;
    int k = numTS - delay;
    k *= (k > 0);
    k %= p;
    k *= (k < length);
    return k;
}

struct ExcEntryLds {
    int comp, x, y, z, delay;
};

// per-cell coefficient row indices of a vector (compressed coefficient modes)
template <int CM, int VZ>
__device__ __forceinline__ void load_rows(const void* idx, long long p, unsigned (&c)[VZ]) {
    if (CM == CF_IDX16) {
        const SVec<unsigned short, VZ> t = *reinterpret_cast<const SVec<unsigned short, VZ>*>((const unsigned short*)idx + p);
#pragma unroll
        for (int k = 0; k < VZ; ++k) c[k] = t.v[k];
    } else {
        const SVec<unsigned, VZ> t = *reinterpret_cast<const SVec<unsigned, VZ>*>((const unsigned*)idx + p);
#pragma unroll
        for (int k = 0; k < VZ; ++k) c[k] = t.v[k];
    }
}

// coefficients of a vector: table rows by prefetched index, or the full arrays
template <typename A, int CM, int VZ>
__device__ __forceinline__ void coeffs_at(const CoeffPtrs<A>& cf, long long p, long long sc, const unsigned (&c)[VZ],
                                          A a[3][VZ], A b[3][VZ]) {
    if (CM == CF_FULL) {
        vload_coeffs<A, CM, VZ>(cf, p, sc, a, b);
    } else {
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

template <typename S, int CM, int VZ, int TY>
__global__ void __launch_bounds__(64 * (TY + 2))
k_step_fused(const S* __restrict__ Vin, const S* __restrict__ Iin, S* __restrict__ Vout, S* __restrict__ Iout,
             CoeffPtrs<typename StoreTraits<S>::A> cfe, CoeffPtrs<typename StoreTraits<S>::A> cfh, Dims d,
             Scales<typename StoreTraits<S>::A> sc, FusedGrid g, ExcDev ev, ExcDev ec, const unsigned* __restrict__ step) {
    typedef StoreTraits<S> T;
    typedef typename T::A A;
    typedef SVec<A, VZ> AV;
    constexpr int TZ = 64 * VZ;
    // only the y-neighbour components go through LDS (z neighbours come from
    // the lane's own registers and wave shuffles); plane-parity double
    // buffers need one barrier per plane
    __shared__ AV sI[2][2][TY + 2][64];        // I_old(x) comps 0 and 2, row w <-> y0-1+w
    __shared__ A sIh[2][2][TY + 2];            // the same at the z-halo cell z0+TZ
    __shared__ AV sV[2][2][TY + 2][64];        // V_new(x) comps 0 and 2, row w (1..TY+1)
    __shared__ ExcEntryLds sEv[FUSED_MAX_EXC], sEc[FUSED_MAX_EXC];
    __shared__ A sAv[FUSED_MAX_EXC], sAc[FUSED_MAX_EXC];

    // XCD-contiguous tile order: block b runs on XCD b % 8
    const int b = blockIdx.x;
    const int t = (b & 7) * g.per_xcd + (b >> 3);
    if (t >= g.nblocks) return;
    const int tz = t % g.tiles_z;
    const int ty = (t / g.tiles_z) % g.tiles_y;
    const int sx = t / (g.tiles_z * g.tiles_y);
    const int y0 = ty * TY, z0 = tz * TZ;
    const int xs = sx * g.seg;
    const int xe = min(xs + g.seg, d.nx);      // planes [xs, xe) owned; xe is the halo plane if < nx

    const int w = threadIdx.x >> 6;            // 0: row above, 1..TY: owned rows, TY+1: halo row
    const int l = threadIdx.x & 63;
    const int y = y0 - 1 + w;
    const bool rowok = y >= 0 && y < d.ny;
    const int zc = z0 + l * VZ;                // first cell of this lane
    const bool zok = zc < d.nz;
    const bool zh_ok = z0 + TZ < d.nz;         // z-halo cell exists (lane 63)
    const int numTS = (int)*step;

    const int nev = min(ev.count, FUSED_MAX_EXC), nec = min(ec.count, FUSED_MAX_EXC);
    for (int i = threadIdx.x; i < nev + nec; i += blockDim.x) {
        const bool cur = i >= nev;
        const int j = cur ? i - nev : i;
        const long long off = cur ? ec.off[j] : ev.off[j];
        const int comp = (int)(off / d.sc);
        const long long cell = off - comp * d.sc;
        ExcEntryLds q;
        q.comp = comp;
        q.x = (int)(cell / d.sx);
        q.y = (int)((cell % d.sx) / d.sy);
        q.z = (int)(cell % d.sy);
        q.delay = cur ? ec.delay[j] : ev.delay[j];
        const A amp = cur ? ((const A*)ec.amp)[j] : ((const A*)ev.amp)[j];
        if (cur) { sEc[j] = q; sAc[j] = amp; } else { sEv[j] = q; sAv[j] = amp; }
    }

    // registers carried along x
    A Ip[3][VZ], Iph[3];                       // I_old(x-1); lane 63 also at z0+TZ
    A Ic[3][VZ], Ich[3], Icm[3];               // I_old(x); lane 63: z0+TZ; lane 0: z0-1
    A Vc[3][VZ], Vch[3];                       // V_old(x); lane 63: z0+TZ
    A Vp[3][VZ], Vph[3];                       // V_new(x-1); lane 63: z0+TZ
    unsigned Rc[VZ], Rp[VZ], Rch;              // coefficient row indices of planes x and x-1; lane 63: z0+TZ
    auto load_plane = [&](int xx, A (&I_)[3][VZ], A (&Ih_)[3], A (&Im_)[3], A (&V_)[3][VZ], A (&Vh_)[3], unsigned (&R_)[VZ],
                          unsigned& Rh_) {
        const bool ok = xx < d.nx && rowok;
        const long long pr = (long long)xx * d.sx + (long long)y * d.sy;
#pragma unroll
        for (int k = 0; k < VZ; ++k) R_[k] = 0;
        Rh_ = 0;
        if (CM != CF_FULL && ok && zok && w >= 1) load_rows<CM, VZ>(cfe.idx, pr + zc, R_);
        if (CM != CF_FULL && ok && l == 63 && zh_ok && w >= 1 && w <= TY)
            Rh_ = CM == CF_IDX16 ? (unsigned)((const unsigned short*)cfe.idx)[pr + z0 + TZ] : ((const unsigned*)cfe.idx)[pr + z0 + TZ];
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            if (ok && zok) {
                vload<S, VZ>(Iin + n * d.sc + pr + zc, I_[n], sc.c[n]);
                if (w >= 1) vload<S, VZ>(Vin + n * d.sc + pr + zc, V_[n], sc.v[n]);
                else {
#pragma unroll
                    for (int k = 0; k < VZ; ++k) V_[n][k] = A(0);
                }
            } else {
#pragma unroll
                for (int k = 0; k < VZ; ++k) { I_[n][k] = A(0); V_[n][k] = A(0); }
            }
            Ih_[n] = (ok && l == 63 && zh_ok) ? T::load(Iin + n * d.sc + pr + z0 + TZ, sc.c[n]) : A(0);
            Vh_[n] = (ok && l == 63 && zh_ok && w >= 1 && w <= TY) ? T::load(Vin + n * d.sc + pr + z0 + TZ, sc.v[n]) : A(0);
            Im_[n] = (ok && l == 0 && z0 > 0) ? T::load(Iin + n * d.sc + pr + z0 - 1, sc.c[n]) : A(0);
        }
    };
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        Iph[n] = A(0); Vph[n] = A(0);
#pragma unroll
        for (int k = 0; k < VZ; ++k) { Ip[n][k] = A(0); Vp[n][k] = A(0); }
    }
    if (xs > 0 && rowok && zok) {
        const long long p = (long long)(xs - 1) * d.sx + (long long)y * d.sy + zc;
#pragma unroll
        for (int n = 0; n < 3; ++n) vload<S, VZ>(Iin + n * d.sc + p, Ip[n], sc.c[n]);
    }
    if (xs > 0 && rowok && l == 63 && zh_ok) {
        const long long p = (long long)(xs - 1) * d.sx + (long long)y * d.sy + z0 + TZ;
#pragma unroll
        for (int n = 0; n < 3; ++n) Iph[n] = T::load(Iin + n * d.sc + p, sc.c[n]);
    }
#pragma unroll
    for (int k = 0; k < VZ; ++k) Rp[k] = 0;
    if (CM != CF_FULL && xs > 0 && rowok && zok && w >= 1)
        load_rows<CM, VZ>(cfe.idx, (long long)(xs - 1) * d.sx + (long long)y * d.sy + zc, Rp);
    load_plane(xs, Ic, Ich, Icm, Vc, Vch, Rc, Rch);
    __syncthreads();   // excitation tables

    const A* sigv = (const A*)ev.sig;
    const A* sigc = (const A*)ec.sig;

    for (int x = xs; x <= xe; ++x) {
        const bool plane = x < d.nx;
        const int par = x & 1;
        const long long prow = (long long)x * d.sx + (long long)y * d.sy;
        // (a) I_old(x) comps 0, 2 of this row for the row below
        {
            AV a0, a2;
#pragma unroll
            for (int k = 0; k < VZ; ++k) { a0.v[k] = Ic[0][k]; a2.v[k] = Ic[2][k]; }
            sI[par][0][w][l] = a0;
            sI[par][1][w][l] = a2;
            if (l == 63) { sIh[par][0][w] = Ich[0]; sIh[par][1][w] = Ich[2]; }
        }
        // z-1 neighbours of the first cell: previous lane's last cell, z0-1 for
        // lane 0.  The shuffles run on every lane (a lane only reads active lanes).
        const A sh0 = __shfl_up(Ic[0][VZ - 1], 1);
        const A sh1 = __shfl_up(Ic[1][VZ - 1], 1);
        const A i0_zm0 = l > 0 ? sh0 : Icm[0];
        const A i1_zm0 = l > 0 ? sh1 : Icm[1];
        __syncthreads();

        // (c) V_new(x) for the owned rows and the halo row
        A Vn[3][VZ], Vnh[3];
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            Vnh[n] = A(0);
#pragma unroll
            for (int k = 0; k < VZ; ++k) Vn[n][k] = A(0);
        }
        if (plane && w >= 1 && rowok && zok) {
            A i0_ym[VZ], i2_ym[VZ];
            if (y > 0) {
                const AV a0 = sI[par][0][w - 1][l], a2 = sI[par][1][w - 1][l];
#pragma unroll
                for (int k = 0; k < VZ; ++k) { i0_ym[k] = a0.v[k]; i2_ym[k] = a2.v[k]; }
            } else {                                    // y-1 neighbour is the cell itself at y = 0
#pragma unroll
                for (int k = 0; k < VZ; ++k) { i0_ym[k] = Ic[0][k]; i2_ym[k] = Ic[2][k]; }
            }
            A curl[3][VZ];
#pragma unroll
            for (int k = 0; k < VZ; ++k) {
                const A i0 = Ic[0][k], i1 = Ic[1][k], i2 = Ic[2][k];
                const A i2_xm = x > 0 ? Ip[2][k] : i2;  // x-1 neighbour, self at x = 0
                const A i1_xm = x > 0 ? Ip[1][k] : i1;
                const A i1_zm = k > 0 ? Ic[1][k - 1] : i1_zm0;
                const A i0_zm = k > 0 ? Ic[0][k - 1] : i0_zm0;
                curl[0][k] = ((i2 - i2_ym[k]) - i1) + i1_zm;
                curl[1][k] = ((i0 - i0_zm) - i2) + i2_xm;
                curl[2][k] = ((i1 - i1_xm) - i0) + i0_ym[k];
            }
            const long long p = prow + zc;
            A vv[3][VZ], vi[3][VZ];
            coeffs_at<A, CM, VZ>(cfe, p, d.sc, Rc, vv, vi);
#pragma unroll
            for (int n = 0; n < 3; ++n)
#pragma unroll
                for (int k = 0; k < VZ; ++k) {
                    A r = Vc[n][k] * vv[n][k];
                    r = r + vi[n][k] * curl[n][k];
                    Vn[n][k] = round_store<S>(r, sc.v[n], sc.iv[n]);
                }
            for (int j = 0; j < nev; ++j) {
                const ExcEntryLds q = sEv[j];
                const int dz = q.z - zc;
                if (q.x == x && q.y == y && dz >= 0 && dz < VZ) {
                    const A s = sigv[exc_index(numTS, q.delay, ev.period, ev.length)];
#pragma unroll
                    for (int k = 0; k < VZ; ++k)
                        if (k == dz)
#pragma unroll
                            for (int n = 0; n < 3; ++n)
                                if (n == q.comp) Vn[n][k] = round_store<S>(Vn[n][k] + sAv[j] * s, sc.v[n], sc.iv[n]);
                }
            }
            if (w <= TY && x < xe) {
                const int valid = min(VZ, d.nz - zc);
#pragma unroll
                for (int n = 0; n < 3; ++n) vstore<S, VZ>(Vout + n * d.sc + p, Vn[n], sc.iv[n], valid);
            }
            // z-halo cell z0+TZ of the owned rows (lane 63), for I_new(x-1, z0+TZ-1)
            if (l == 63 && w <= TY && zh_ok) {
                const long long ph = prow + z0 + TZ;
                const A i0 = Ich[0], i1 = Ich[1], i2 = Ich[2];
                const A i0_ymh = y > 0 ? sIh[par][0][w - 1] : i0;
                const A i2_ymh = y > 0 ? sIh[par][1][w - 1] : i2;
                const A i2_xm = x > 0 ? Iph[2] : i2;
                const A i1_xm = x > 0 ? Iph[1] : i1;
                const A i1_zm = Ic[1][VZ - 1], i0_zm = Ic[0][VZ - 1];
                A ch[3];
                ch[0] = ((i2 - i2_ymh) - i1) + i1_zm;
                ch[1] = ((i0 - i0_zm) - i2) + i2_xm;
                ch[2] = ((i1 - i1_xm) - i0) + i0_ymh;
                A vv3[3], vi3[3];
                if (CM == CF_FULL) {
                    load_coeffs<A, CM>(cfe, ph, d.sc, vv3, vi3);
                } else {
                    const A* row = cfe.a + (size_t)Rch * 8;
#pragma unroll
                    for (int n = 0; n < 3; ++n) { vv3[n] = row[n]; vi3[n] = row[4 + n]; }
                }
#pragma unroll
                for (int n = 0; n < 3; ++n) {
                    A r = Vch[n] * vv3[n];
                    r = r + vi3[n] * ch[n];
                    Vnh[n] = round_store<S>(r, sc.v[n], sc.iv[n]);
                }
                for (int j = 0; j < nev; ++j) {
                    const ExcEntryLds q = sEv[j];
                    if (q.x == x && q.y == y && q.z == z0 + TZ) {
                        const A s = sigv[exc_index(numTS, q.delay, ev.period, ev.length)];
#pragma unroll
                        for (int n = 0; n < 3; ++n)
                            if (n == q.comp) Vnh[n] = round_store<S>(Vnh[n] + sAv[j] * s, sc.v[n], sc.iv[n]);
                    }
                }
            }
        }
        if (w >= 1) {
            AV a0, a2;
#pragma unroll
            for (int k = 0; k < VZ; ++k) { a0.v[k] = Vn[0][k]; a2.v[k] = Vn[2][k]; }
            sV[par][0][w][l] = a0;
            sV[par][1][w][l] = a2;
        }
        // the next plane's loads go out before the current update
        A In[3][VZ], Inh[3], Inm[3], Vx[3][VZ], Vxh[3];
        unsigned Rn[VZ], Rnh;
        load_plane(x + 1 <= xe ? x + 1 : d.nx, In, Inh, Inm, Vx, Vxh, Rn, Rnh);

        // (e) I_new(x-1) for the owned rows.  V_new(x-1, y+1) is in sV of the
        // previous plane's parity, written before this plane's barrier.
        const int xi = x - 1;
        // z+1 neighbour of the last cell: next lane's first cell, the z-halo for lane 63
        const A sd0 = __shfl_down(Vp[0][0], 1);
        const A sd1 = __shfl_down(Vp[1][0], 1);
        const A v0_zpl = l < 63 ? sd0 : Vph[0];
        const A v1_zpl = l < 63 ? sd1 : Vph[1];
        if (xi >= xs && w >= 1 && w <= TY && rowok && zok) {
            const long long p = (long long)xi * d.sx + (long long)y * d.sy + zc;
            A c[3][VZ];
            if (xi <= d.nx - 2) {
                const AV b0 = sV[par ^ 1][0][w + 1][l], b2 = sV[par ^ 1][1][w + 1][l];
                A curl[3][VZ];
#pragma unroll
                for (int k = 0; k < VZ; ++k) {
                    const A v0 = Vp[0][k], v1 = Vp[1][k], v2 = Vp[2][k];
                    const A v2_yp = b2.v[k], v0_yp = b0.v[k];
                    const A v1_zp = k < VZ - 1 ? Vp[1][k + 1] : v1_zpl;
                    const A v0_zp = k < VZ - 1 ? Vp[0][k + 1] : v0_zpl;
                    const A v2_xp = Vn[2][k], v1_xp = Vn[1][k];
                    curl[0][k] = ((v2 - v2_yp) - v1) + v1_zp;
                    curl[1][k] = ((v0 - v0_zp) - v2) + v2_xp;
                    curl[2][k] = ((v1 - v1_xp) - v0) + v0_yp;
                }
                A ii[3][VZ], iv[3][VZ];
                coeffs_at<A, CM, VZ>(cfh, p, d.sc, Rp, ii, iv);
                const bool yup = y < d.ny - 1;
#pragma unroll
                for (int n = 0; n < 3; ++n)
#pragma unroll
                    for (int k = 0; k < VZ; ++k) {
                        if (yup && zc + k < d.nz - 1) {
                            A r = Ip[n][k] * ii[n][k];
                            r = r + iv[n][k] * curl[n][k];
                            c[n][k] = round_store<S>(r, sc.c[n], sc.ic[n]);
                        } else {
                            c[n][k] = Ip[n][k];          // not updated: carried over
                        }
                    }
            } else {
#pragma unroll
                for (int n = 0; n < 3; ++n)
#pragma unroll
                    for (int k = 0; k < VZ; ++k) c[n][k] = Ip[n][k];   // plane Nx-1: carried over
            }
            for (int j = 0; j < nec; ++j) {
                const ExcEntryLds q = sEc[j];
                const int dz = q.z - zc;
                if (q.x == xi && q.y == y && dz >= 0 && dz < VZ) {
                    const A s = sigc[exc_index(numTS, q.delay, ec.period, ec.length)];
#pragma unroll
                    for (int k = 0; k < VZ; ++k)
                        if (k == dz)
#pragma unroll
                            for (int n = 0; n < 3; ++n)
                                if (n == q.comp) c[n][k] = round_store<S>(c[n][k] + sAc[j] * s, sc.c[n], sc.ic[n]);
                }
            }
            const int valid = min(VZ, d.nz - zc);
#pragma unroll
            for (int n = 0; n < 3; ++n) vstore<S, VZ>(Iout + n * d.sc + p, c[n], sc.ic[n], valid);
        }
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            Iph[n] = Ich[n];
            Ich[n] = Inh[n];
            Icm[n] = Inm[n];
            Vch[n] = Vxh[n];
            Vph[n] = Vnh[n];
#pragma unroll
            for (int k = 0; k < VZ; ++k) {
                Ip[n][k] = Ic[n][k]; Ic[n][k] = In[n][k]; Vc[n][k] = Vx[n][k]; Vp[n][k] = Vn[n][k];
            }
        }
#pragma unroll
        for (int k = 0; k < VZ; ++k) { Rp[k] = Rc[k]; Rc[k] = Rn[k]; }
        Rch = Rnh;
    }
}
