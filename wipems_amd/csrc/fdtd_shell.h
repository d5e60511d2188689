// Partition of the vector kernels' work between an interior launch without
// UPML code and a launch over the UPML shell (host side; shared by the
// engine and its host-only partition check).
#pragma once
#include "fdtd_device.h"

// The interior box left free by the slabs: on each axis, the planes the
// slabs cover completely at either end are outside it (slabs are pairwise
// disjoint, as fdtd_add_upml checks, so a plane is covered when the areas of
// the slabs crossing it add up to the plane's).  The slabs form its shell
// when none meets the box and their cells add up to the domain minus the
// box.  Returns false (ibox untouched) when they do not.
inline bool find_interior(const uint32_t N[3], const PmlSlab* slab, int count, int ibox[3][2]) {
    if (count <= 0) return false;
    int lo[3], hi[3];
    for (int n = 0; n < 3; ++n) {
        const int a = (n + 1) % 3, b = (n + 2) % 3;
        auto covered = [&](int c) {
            long long area = 0;
            for (int s = 0; s < count; ++s)
                if (c >= slab[s].s[n] && c < slab[s].s[n] + slab[s].n[n]) area += (long long)slab[s].n[a] * slab[s].n[b];
            return area == (long long)N[a] * N[b];
        };
        lo[n] = 0;
        while (lo[n] < (int)N[n] && covered(lo[n])) ++lo[n];
        hi[n] = (int)N[n];
        while (hi[n] > lo[n] && covered(hi[n] - 1)) --hi[n];
    }
    long long box = 1;
    for (int n = 0; n < 3; ++n) box *= std::max(0, hi[n] - lo[n]);
    long long cells = 0;
    for (int s = 0; s < count; ++s) {
        const PmlSlab& b = slab[s];
        bool meets = box > 0;
        for (int n = 0; n < 3; ++n)
            if (b.s[n] + b.n[n] <= lo[n] || b.s[n] >= hi[n]) meets = false;
        if (meets) return false;
        cells += (long long)b.n[0] * b.n[1] * b.n[2];
    }
    if (cells + box != (long long)N[0] * N[1] * N[2]) return false;
    for (int n = 0; n < 3; ++n) {
        ibox[n][0] = lo[n];
        ibox[n][1] = std::max(lo[n], hi[n]);
    }
    return true;
}

// Launch geometry of one vector-kernel half step.  Without UPML (or with
// slabs that do not form a shell) one launch covers planes [0, xe) and rows
// [0, ye); with a shell, the main launch covers the vectors of the interior
// box only and the shell launch the rest.
struct VecLaunch {
    VecRange r;
    unsigned nx;               // planes of the main launch (0: none)
    long long nq;              // flat (y, z) extent of the main launch per plane
    bool pml_main;             // main launch runs the UPML form
    bool shell;
    ShellGrid g;
};

// za: the interior launch's z range starts and ends at multiples of za
// vectors (rounded inwards), so that it and the shell launch share fewer
// partial cache lines at the ends of the interior rows.
inline VecLaunch plan_vec_launch(const uint32_t N[3], long long sy, int pml_count, bool shell_ok, const int ibox[3][2], int vz, bool curr,
                                 int za = 1) {
    VecLaunch L{};
    const int nx = (int)N[0], ny = (int)N[1], nz = (int)N[2];
    const int xe = curr ? std::max(1, nx - 1) : nx;
    const int ye = curr ? ny - 1 : ny;
    L.pml_main = pml_count > 0 && !shell_ok;
    L.shell = pml_count > 0 && shell_ok;
    if (!L.shell) {
        L.r = VecRange{0, 0, ny, 0, 1 << 30, 0, 0, 0, 0};
        L.nx = (unsigned)xe;
        L.nq = (long long)std::max(ye, 0) * sy;
        return L;
    }
    const int x0 = std::min(ibox[0][0], xe), x1 = std::min(std::max(ibox[0][1], x0), xe);
    const int y0 = ibox[1][0], y1 = std::max(ibox[1][1], y0);
    const int vn = (nz + vz - 1) / vz;
    za = std::max(za, 1);
    const int vlo = std::min(((ibox[2][0] + vz - 1) / vz + za - 1) / za * za, vn);
    const int vhi = std::max(std::min(ibox[2][1] / vz / za * za, vn), vlo);
    L.r = VecRange{x0, y0, y1, vlo * vz, vhi * vz, 0, 0, 0, 0};
    L.nx = (unsigned)(x1 - x0);
    L.nq = (long long)(y1 - y0) * sy;
    if (y1 == y0 || vhi == vlo) L.nx = 0;
    ShellGrid& g = L.g;
    g.x0 = x0; g.x1 = x1; g.xe = xe;
    g.y0 = y0; g.y1 = y1; g.ny = ny;
    g.vlo = vlo; g.vhi = vhi; g.vn = vn;
    g.nA0 = (long long)x0 * ny * vn;
    g.nA1 = (long long)(xe - x1) * ny * vn;
    g.nB = (long long)(x1 - x0) * (y0 + ny - y1) * vn;
    g.nC = (long long)(x1 - x0) * (y1 - y0) * (vlo + vn - vhi);
    return L;
}

// The shell of L (planned with vectors of vz cells) in vectors of vz / f
// cells: the interior range [vlo*vz, vhi*vz) stays the same, so the two
// launches still partition the vectors' cells.
inline ShellGrid shell_grid_split(const VecLaunch& L, int nz, int vz, int f) {
    ShellGrid g = L.g;
    const int w = vz / f;
    g.vn = (nz + w - 1) / w;
    g.vlo = std::min(L.g.vlo * f, g.vn);
    g.vhi = std::max(std::min(L.g.vhi * f, g.vn), g.vlo);
    g.nA0 = (long long)g.x0 * g.ny * g.vn;
    g.nA1 = (long long)(g.xe - g.x1) * g.ny * g.vn;
    g.nB = (long long)(g.x1 - g.x0) * (g.y0 + g.ny - g.y1) * g.vn;
    g.nC = (long long)(g.x1 - g.x0) * (g.y1 - g.y0) * (g.vlo + g.vn - g.vhi);
    return g;
}
