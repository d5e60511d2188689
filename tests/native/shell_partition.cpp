// Host-only check of the split vector launches (fdtd_shell.h): for UPML slab
// sets read from stdin, every vector the update bodies process is covered
// exactly once by the interior launch plus the shell launch, the interior
// launch touches no slab cell, and the shell enumeration stays inside the
// field arrays.  Input: cases of "N0 N1 N2 count" followed by count lines
// "s0 s1 s2 n0 n1 n2".  Prints one line per case and vector width; exit 1
// on the first violation.
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "fdtd_shell.h"

static int check(const uint32_t N[3], const std::vector<PmlSlab>& slabs, int vz, bool curr) {
    int ibox[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    const bool ok = find_interior(N, slabs.data(), (int)slabs.size(), ibox);
    const long long sy = ((long long)N[2] + 15) / 16 * 16;
    const VecLaunch L = plan_vec_launch(N, sy, (int)slabs.size(), ok, ibox, vz, curr);
    const int nx = (int)N[0], ny = (int)N[1], nz = (int)N[2];
    const int vrow = (int)(sy / vz);
    std::vector<unsigned char> cnt((size_t)nx * ny * vrow, 0);
    // cells the update body processes: volt y < ny, z0 < nz; curr x < nx-1, y < ny-1, z0 < nz-1
    auto body = [&](int x, int y, int z0) {
        if (curr) return x < std::max(1, nx - 1) && y < ny - 1 && z0 < nz - 1;
        return x < nx && y < ny && z0 < nz;
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
            if (!body(x, y, z0)) continue;
            if (L.shell && in_slab(x, y, z0)) { printf("interior launch reaches a slab cell at %d %d %d\n", x, y, z0); return 1; }
            ++cnt[((size_t)x * ny + y) * vrow + z0 / vz];
            ++main_vecs;
        }
    }
    if (L.shell) {
        const long long total = L.g.nA0 + L.g.nA1 + L.g.nB + L.g.nC;
        for (long long t = 0; t < (total + 255) / 256 * 256; ++t) {
            int x;
            long long q0;
            if (!shell_vector(L.g, t, vz, sy, x, q0)) {
                if (t < total) { printf("shell thread %lld of %lld has no vector\n", t, total); return 1; }
                continue;
            }
            if (t >= total || x < 0 || x >= nx || q0 < 0 || q0 + vz > (long long)ny * sy || q0 % vz) {
                printf("shell thread %lld maps outside the fields: x %d q0 %lld\n", t, x, q0);
                return 1;
            }
            const int y = (int)(q0 / sy), z0 = (int)(q0 - (long long)y * sy);
            if (!body(x, y, z0)) continue;
            ++cnt[((size_t)x * ny + y) * vrow + z0 / vz];
            ++shell_vecs;
        }
    }
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int v = 0; v < vrow; ++v) {
                const int c = cnt[((size_t)x * ny + y) * vrow + v];
                const int want = body(x, y, v * vz) ? 1 : 0;
                if (c != want) {
                    printf("vector %d %d %d (vz %d, %s) covered %d times\n", x, y, v * vz, vz, curr ? "curr" : "volt", c);
                    return 1;
                }
            }
    printf("N %u %u %u slabs %zu shell %d vz %d %s: interior %lld shell %lld\n", N[0], N[1], N[2], slabs.size(), (int)ok, vz,
           curr ? "curr" : "volt", main_vecs, shell_vecs);
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
            for (bool curr : {false, true})
                if (check(N, slabs, vz, curr)) return 1;
    }
    return 0;
}
