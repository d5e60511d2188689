// Host-only check of the split vector launches (fdtd_shell.h): for UPML slab
// sets read from stdin, every cell the update bodies write is covered
// exactly once by the interior launch plus the shell launch (the shell in
// vectors of the interior's width, half or a quarter of it; the interior's z
// range aligned to 1, 4 or 8 vectors), the interior launch
// touches no slab cell, and the shell enumeration stays inside the field
// arrays.  Input: cases of "N0 N1 N2 count" followed by count lines
// "s0 s1 s2 n0 n1 n2".  Prints one line per case and vector width; exit 1
// on the first violation.
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "fdtd_shell.h"

static int check(const uint32_t N[3], const std::vector<PmlSlab>& slabs, int vz, bool curr, int f, int za) {
    int ibox[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    const bool ok = find_interior(N, slabs.data(), (int)slabs.size(), ibox);
    const long long sy = ((long long)N[2] + 15) / 16 * 16;
    const VecLaunch L = plan_vec_launch(N, sy, (int)slabs.size(), ok, ibox, vz, curr, za);
    const ShellGrid G = f == 1 ? L.g : shell_grid_split(L, (int)N[2], vz, f);
    const int w = vz / f;
    const int nx = (int)N[0], ny = (int)N[1], nz = (int)N[2];
    // cells the update bodies write: volt x < nx, y < ny, z < nz; curr
    // x < max(1, nx-1), y < ny-1, z < nz-1 (a vector starting beyond them
    // returns, one reaching past them stores its valid cells only)
    const int xe = curr ? std::max(1, nx - 1) : nx, ye = curr ? ny - 1 : ny, ze = curr ? nz - 1 : nz;
    std::vector<unsigned char> cnt((size_t)nx * ny * nz, 0);
    auto cover = [&](int x, int y, int z0, int width) {
        if (x >= xe || y >= ye || z0 >= ze) return 0;
        for (int z = z0; z < std::min(z0 + width, ze); ++z) ++cnt[((size_t)x * ny + y) * nz + z];
        return 1;
    };
    auto in_slab = [&](int x, int y, int z0) {
        for (const PmlSlab& b : slabs)
            for (int k = 0; k < vz; ++k)
                if (x >= b.s[0] && x < b.s[0] + b.n[0] && y >= b.s[1] && y < b.s[1] + b.n[1] &&
                    z0 + k >= b.s[2] && z0 + k < b.s[2] + b.n[2]) return true;
        return false;
    };
    long long main_vecs = 0, shell_vecs = 0;
    for (unsigned bx = 0; bx < L.nx; ++bx) {
        const long long threads = (L.nq / vz + 255) / 256 * 256;
        for (long long t = 0; t < threads; ++t) {
            const long long q0 = (long long)L.r.y0 * sy + t * vz;
            const int y = (int)(q0 / sy), z0 = (int)(q0 - (long long)y * sy);
            if (y >= L.r.y1 || z0 < L.r.zlo || z0 + vz > L.r.zhi) continue;
            const int x = L.r.x0 + (int)bx;
            if (x >= xe || y >= ye || z0 >= ze) continue;
            if (L.shell && in_slab(x, y, z0)) { printf("interior launch reaches a slab cell at %d %d %d\n", x, y, z0); return 1; }
            main_vecs += cover(x, y, z0, vz);
        }
    }
    if (L.shell) {
        const long long total = G.nA0 + G.nA1 + G.nB + G.nC;
        for (long long t = 0; t < (total + 255) / 256 * 256; ++t) {
            int x;
            long long q0;
            if (!shell_vector(G, t, w, sy, x, q0)) {
                if (t < total) { printf("shell thread %lld of %lld has no vector\n", t, total); return 1; }
                continue;
            }
            if (t >= total || x < 0 || x >= nx || q0 < 0 || q0 + w > (long long)ny * sy || q0 % w) {
                printf("shell thread %lld maps outside the fields: x %d q0 %lld\n", t, x, q0);
                return 1;
            }
            const int y = (int)(q0 / sy), z0 = (int)(q0 - (long long)y * sy);
            shell_vecs += cover(x, y, z0, w);
        }
    }
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                const int c = cnt[((size_t)x * ny + y) * nz + z];
                const int want = x < xe && y < ye && z < ze ? 1 : 0;
                if (c != want) {
                    printf("cell %d %d %d (vz %d, shell vectors of %d, %s) covered %d times\n", x, y, z, vz, w, curr ? "curr" : "volt", c);
                    return 1;
                }
            }
    printf("N %u %u %u slabs %zu shell %d vz %d/%d za %d %s: interior %lld shell %lld\n", N[0], N[1], N[2], slabs.size(), (int)ok, vz, w,
           za, curr ? "curr" : "volt", main_vecs, shell_vecs);
    return 0;
}

int main() {
    uint32_t N[3];
    int count;
    while (scanf("%u %u %u %d", &N[0], &N[1], &N[2], &count) == 4) {
        std::vector<PmlSlab> slabs(count);
        for (int s = 0; s < count; ++s) {
            PmlSlab b{};
            if (scanf("%d %d %d %d %d %d", &b.s[0], &b.s[1], &b.s[2], &b.n[0], &b.n[1], &b.n[2]) != 6) return 2;
            slabs[s] = b;
        }
        for (int vz : {2, 4, 8})
            for (int f : {1, 2, 4})
                for (int za : {1, 4, 8})
                    for (bool curr : {false, true})
                        if (f <= vz && check(N, slabs, vz, curr, f, za)) return 1;
    }
    return 0;
}
