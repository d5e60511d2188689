// nf2ff.hip -- near-field to far-field transform (include/wipems_nf2ff.h).
//
// nf2ff_calc::AddSinglePlane (nf2ff/nf2ff_calc.cpp:357-528) and its worker
// (:40-163) for a Cartesian face without mirrors.  The host forms the
// area-weighted surface currents a*Js = a*(n x H) and a*Ms = -a*(n x E) per
// node (six complex values); k_nf2ff_face gives each (theta, phi) direction
// one workgroup, which sums those currents times exp(jk (r - c).rhat) over
// all nodes in FP64; the host turns the six sums into N_theta, N_phi, L_theta
// and L_phi and accumulates the far fields in the reference's formulas.
#include "fdtd_internal.h"
#include "wipems_nf2ff.h"

#include <complex>
#include <cstring>
#include <memory>
#include <vector>

namespace {

const double kC0 = 299792458.0;                     // __C0__, tools/constants.h:25
const double kZ0 = 376.730313461;                   // __Z0__, tools/constants.h:26

struct FaceDev {
    int n[3];                   // node counts, one of them 1
    const double* line[3];      // node coordinates (m), relative to the centre
    const double2* cur;         // [6][nodes]: a*Js x, y, z, a*Ms x, y, z
    long long nodes;
    const double* dir;          // [n_dir][3]: rhat
    double k;
    double2* out;               // [n_dir][6] sums
};

__global__ void __launch_bounds__(256) k_nf2ff_face(FaceDev f) {
    const int a = blockIdx.x;
    const double ux = f.dir[3 * a], uy = f.dir[3 * a + 1], uz = f.dir[3 * a + 2];
    double2 s[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) s[q] = make_double2(0.0, 0.0);
    const long long n12 = (long long)f.n[1] * f.n[2];
    for (long long i = threadIdx.x; i < f.nodes; i += blockDim.x) {
        const int x = (int)(i / n12);
        const long long r = i - (long long)x * n12;
        const int y = (int)(r / f.n[2]);
        const int z = (int)(r - (long long)y * f.n[2]);
        const double arg = f.k * (f.line[0][x] * ux + f.line[1][y] * uy + f.line[2][z] * uz);
        double sn, cs;
        sincos(arg, &sn, &cs);
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const double2 c = f.cur[q * f.nodes + i];
            s[q].x += c.x * cs - c.y * sn;
            s[q].y += c.x * sn + c.y * cs;
        }
    }
    __shared__ double red[256];
    for (int q = 0; q < 6; ++q) {
        for (int part = 0; part < 2; ++part) {
            red[threadIdx.x] = part ? s[q].y : s[q].x;
            __syncthreads();
            for (int w = 128; w > 0; w >>= 1) {
                if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
                __syncthreads();
            }
            if (threadIdx.x == 0) {
                if (part) f.out[6 * a + q].y = red[0];
                else f.out[6 * a + q].x = red[0];
            }
            __syncthreads();
        }
    }
}

std::vector<double> half_widths(const double* l, uint32_t n) {
    std::vector<double> w(n, 1.0);
    if (n < 2) return w;
    for (uint32_t i = 1; i + 1 < n; ++i) w[i] = 0.5 * std::fabs(l[i + 1] - l[i - 1]);
    w[0] = 0.5 * std::fabs(l[1] - l[0]);
    w[n - 1] = 0.5 * std::fabs(l[n - 1] - l[n - 2]);
    return w;
}

}  // namespace

struct nf2ff_calc {
    double freq = 0, k = 0, Z = 0, radius = 1;
    double center[3] = {0, 0, 0};
    std::vector<double> theta, phi;
    std::vector<std::complex<double>> Et, Ep, Ht, Hp;
    std::vector<double> Prad;
    double rad_power = 0, max_dir = 0;
    int mirror[3] = {0, 0, 0};      // 0 off, 1 PEC, 2 PMC
    double mirror_pos[3] = {0, 0, 0};
    hipStream_t stream = nullptr;
    DevBuf dir, out;
    ~nf2ff_calc() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

extern "C" {

int nf2ff_create(double freq, const double* theta, uint32_t n_theta, const double* phi, uint32_t n_phi,
                 const double center[3], double radius, double eps_r, double mue_r, nf2ff_calc** out) {
    if (!theta || !phi || !center || !out || !n_theta || !n_phi) return fail(FDTD_ERR_ARG, "bad argument");
    if (!(freq > 0) || !(radius > 0) || !(eps_r > 0) || !(mue_r > 0)) return fail(FDTD_ERR_ARG, "frequency, radius, eps and mu must be positive");
    if ((uint64_t)n_theta * n_phi > 0x7fffffffULL) return fail(FDTD_ERR_RANGE, "too many directions");
    auto c = std::unique_ptr<nf2ff_calc>(new nf2ff_calc());
    c->freq = freq;
    c->k = 2 * M_PI * freq / kC0 * std::sqrt(eps_r * mue_r);
    c->Z = kZ0 * std::sqrt(mue_r / eps_r);
    c->radius = radius;
    for (int n = 0; n < 3; ++n) c->center[n] = center[n];
    c->theta.assign(theta, theta + n_theta);
    c->phi.assign(phi, phi + n_phi);
    const size_t nd = (size_t)n_theta * n_phi;
    c->Et.assign(nd, 0); c->Ep.assign(nd, 0); c->Ht.assign(nd, 0); c->Hp.assign(nd, 0);
    c->Prad.assign(nd, 0.0);
    std::vector<double> u(3 * nd);
    for (uint32_t t = 0; t < n_theta; ++t)
        for (uint32_t p = 0; p < n_phi; ++p) {
            const size_t a = (size_t)t * n_phi + p;
            u[3 * a] = std::cos(phi[p]) * std::sin(theta[t]);
            u[3 * a + 1] = std::sin(phi[p]) * std::sin(theta[t]);
            u[3 * a + 2] = std::cos(theta[t]);
        }
    int rc;
    if ((rc = c->dir.alloc(u.size() * sizeof(double))) || (rc = c->out.alloc(nd * 6 * sizeof(double2)))) return rc;
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(hipMemcpyAsync(c->dir.p, u.data(), u.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *out = c.release();
    return FDTD_OK;
}

}  // extern "C"

// nf2ff_calc::AddSinglePlane for one face (lines in metres, fields
// [3][nx][ny][nz] complex)
static int add_single(nf2ff_calc* c, const double* const lines[3], const uint32_t n[3], const double* E, const double* H) {
    const uint32_t nx = n[0], ny = n[1], nz = n[2];
    // nf2ff_calc::GetNormalDir (:216-229): the last axis with one line whose
    // two others have more than two
    int nrm = -1;
    for (int a = 0; a < 3; ++a)
        if (n[a] == 1 && n[(a + 1) % 3] > 2 && n[(a + 2) % 3] > 2) nrm = a;
    if (nrm < 0) return fail(FDTD_ERR_ARG, "cannot determine the normal direction of the plane");
    const int nP = (nrm + 1) % 3, nPP = (nrm + 2) % 3;
    const long long nodes = (long long)nx * ny * nz;
    const double sign = lines[nrm][0] >= c->center[nrm] ? 1.0 : -1.0;
    double nd[3] = {0, 0, 0};
    nd[nrm] = sign;
    const std::vector<double> wP = half_widths(lines[nP], n[nP]), wPP = half_widths(lines[nPP], n[nPP]);
    typedef std::complex<double> cd;
    const cd* Ec = reinterpret_cast<const cd*>(E);
    const cd* Hc = reinterpret_cast<const cd*>(H);
    std::vector<cd> cur(6 * (size_t)nodes);
    double power = 0;
    for (uint32_t x = 0; x < nx; ++x)
        for (uint32_t y = 0; y < ny; ++y)
            for (uint32_t z = 0; z < nz; ++z) {
                const size_t i = ((size_t)x * ny + y) * nz + z;
                const uint32_t pos[3] = {x, y, z};
                const double area = wP[pos[nP]] * wPP[pos[nPP]];
                const cd e[3] = {Ec[i], Ec[nodes + i], Ec[2 * nodes + i]};
                const cd h[3] = {Hc[i], Hc[nodes + i], Hc[2 * nodes + i]};
                // radiated power through the face (:414-423)
                power += 0.5 * area * (e[nP] * std::conj(h[nPP]) - e[nPP] * std::conj(h[nP])).real() * sign;
                // Js = n x H, Ms = -n x E (:66-78)
                cur[i] = area * (nd[1] * h[2] - nd[2] * h[1]);
                cur[nodes + i] = area * (nd[2] * h[0] - nd[0] * h[2]);
                cur[2 * nodes + i] = area * (nd[0] * h[1] - nd[1] * h[0]);
                cur[3 * nodes + i] = area * (nd[2] * e[1] - nd[1] * e[2]);
                cur[4 * nodes + i] = area * (nd[0] * e[2] - nd[2] * e[0]);
                cur[5 * nodes + i] = area * (nd[1] * e[0] - nd[0] * e[1]);
            }
    c->rad_power += power;
    DevBuf dl, dc;
    int rc;
    if ((rc = dl.alloc((size_t)(nx + ny + nz) * sizeof(double))) || (rc = dc.alloc(cur.size() * sizeof(double2)))) return rc;
    std::vector<double> rel;
    rel.reserve(nx + ny + nz);
    for (int a = 0; a < 3; ++a)
        for (uint32_t i = 0; i < n[a]; ++i) rel.push_back(lines[a][i] - c->center[a]);
    HIPCHK(hipMemcpyAsync(dl.p, rel.data(), rel.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(dc.p, cur.data(), cur.size() * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    FaceDev f;
    for (int a = 0; a < 3; ++a) f.n[a] = (int)n[a];
    f.line[0] = dl.as<double>();
    f.line[1] = f.line[0] + nx;
    f.line[2] = f.line[1] + ny;
    f.cur = dc.as<double2>();
    f.nodes = nodes;
    f.dir = c->dir.as<double>();
    f.k = c->k;
    f.out = c->out.as<double2>();
    const size_t nd_ = c->theta.size() * c->phi.size();
    hipLaunchKernelGGL(k_nf2ff_face, dim3((unsigned)nd_), dim3(256), 0, c->stream, f);
    HIPCHK(hipGetLastError());
    std::vector<cd> sums(6 * nd_);
    HIPCHK(hipMemcpyAsync(sums.data(), c->out.p, sums.size() * sizeof(cd), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    // equations 8.23a/b, 8.24a/b (:488-505)
    const cd factor = cd(0.0, c->k / 4.0 / M_PI / c->radius) * std::exp(cd(0.0, -c->k * c->radius));
    double pmax = 0;
    const size_t np = c->phi.size();
    for (size_t t = 0; t < c->theta.size(); ++t) {
        const double st = std::sin(c->theta[t]), ct = std::cos(c->theta[t]);
        for (size_t p = 0; p < np; ++p) {
            const size_t a = t * np + p;
            const double sp = std::sin(c->phi[p]), cp = std::cos(c->phi[p]);
            const cd* S = &sums[6 * a];
            const cd Nt = S[0] * (ct * cp) + S[1] * (ct * sp) - S[2] * st;
            const cd Np = S[1] * cp - S[0] * sp;
            const cd Lt = S[3] * (ct * cp) + S[4] * (ct * sp) - S[5] * st;
            const cd Lp = S[4] * cp - S[3] * sp;
            c->Et[a] -= factor * (Lp + c->Z * Nt);
            c->Ep[a] += factor * (Lt - c->Z * Np);
            c->Ht[a] += factor * (Np - Lt / c->Z);
            c->Hp[a] -= factor * (Nt + Lp / c->Z);
            c->Prad[a] = (std::norm(c->Et[a]) + std::norm(c->Ep[a])) / (2 * c->Z);
            pmax = std::max(pmax, c->Prad[a]);
        }
    }
    c->max_dir = 4 * M_PI * pmax / c->rad_power;
    return FDTD_OK;
}

// AddMirrorPlane (:247-279): mirror axis a's lines about the plane, scale the
// fields (PEC: normal H and tangential E flip; PMC: normal E and tangential
// H flip) in place, add the image face
static int add_mirror(nf2ff_calc* c, int a, std::vector<double> (&lines)[3], const uint32_t n[3], std::vector<double>& E,
                      std::vector<double>& H) {
    for (auto& v : lines[a]) v = 2.0 * c->mirror_pos[a] - v;
    double ef[3] = {1, 1, 1}, hf[3] = {1, 1, 1};
    const int aP = (a + 1) % 3, aPP = (a + 2) % 3;
    if (c->mirror[a] == 1) { hf[a] = -1; ef[aP] = ef[aPP] = -1; }
    else { ef[a] = -1; hf[aP] = hf[aPP] = -1; }
    const size_t nodes = (size_t)n[0] * n[1] * n[2];
    for (int d = 0; d < 3; ++d)
        for (size_t i = 0; i < 2 * nodes; ++i) {
            E[2 * d * nodes + i] *= ef[d];
            H[2 * d * nodes + i] *= hf[d];
        }
    const double* lp[3] = {lines[0].data(), lines[1].data(), lines[2].data()};
    return add_single(c, lp, n, E.data(), H.data());
}

extern "C" {

int nf2ff_set_mirror(nf2ff_calc* c, int type, int axis, double pos) {
    if (!c) return fail(FDTD_ERR_ARG, "null handle");
    if (axis < 0 || axis > 2) return fail(FDTD_ERR_RANGE, "mirror axis must be 0..2");
    if (type < 0 || type > 2) return fail(FDTD_ERR_ARG, "mirror type must be 0 (off), 1 (PEC) or 2 (PMC)");
    c->mirror[axis] = type;
    c->mirror_pos[axis] = pos;
    return FDTD_OK;
}

// AddPlane (:281-328): the face, then its images in the reference's order
int nf2ff_add_plane(nf2ff_calc* c, const double* lines_x, uint32_t nx, const double* lines_y, uint32_t ny,
                    const double* lines_z, uint32_t nz, const double* E, const double* H) {
    if (!c || !lines_x || !lines_y || !lines_z || !E || !H) return fail(FDTD_ERR_ARG, "null argument");
    const uint32_t n[3] = {nx, ny, nz};
    const double* lp[3] = {lines_x, lines_y, lines_z};
    int rc;
    if ((rc = add_single(c, lp, n, E, H))) return rc;
    const int* m = c->mirror;
    if (!m[0] && !m[1] && !m[2]) return FDTD_OK;
    const size_t vals = 6 * (size_t)nx * ny * nz;
    std::vector<double> lines[3] = {std::vector<double>(lines_x, lines_x + nx), std::vector<double>(lines_y, lines_y + ny),
                                    std::vector<double>(lines_z, lines_z + nz)};
    std::vector<double> e(E, E + vals), h(H, H + vals);
    for (int a = 0; a < 3; ++a) {
        const int aP = (a + 1) % 3, aPP = (a + 2) % 3;
        if (m[a] && !m[aP] && !m[aPP]) return add_mirror(c, a, lines, n, e, h);
        if (!m[a] && m[aP] && m[aPP]) {
            if ((rc = add_mirror(c, aP, lines, n, e, h)) || (rc = add_mirror(c, aPP, lines, n, e, h))) return rc;
            return add_mirror(c, aP, lines, n, e, h);
        }
    }
    if (m[0] && m[1] && m[2])
        for (int a : {0, 1, 0, 2, 0, 1, 0})
            if ((rc = add_mirror(c, a, lines, n, e, h))) return rc;
    return FDTD_OK;
}

int nf2ff_get(nf2ff_calc* c, double* E_theta, double* E_phi, double* H_theta, double* H_phi, double* P_rad,
              double* rad_power, double* max_dir) {
    if (!c) return fail(FDTD_ERR_ARG, "null handle");
    const size_t nd = c->Et.size();
    auto put = [&](const std::vector<std::complex<double>>& v, double* o) {
        if (o) memcpy(o, v.data(), nd * sizeof(std::complex<double>));
    };
    put(c->Et, E_theta);
    put(c->Ep, E_phi);
    put(c->Ht, H_theta);
    put(c->Hp, H_phi);
    if (P_rad) memcpy(P_rad, c->Prad.data(), nd * sizeof(double));
    if (rad_power) *rad_power = c->rad_power;
    if (max_dir) *max_dir = c->max_dir;
    return FDTD_OK;
}

int nf2ff_destroy(nf2ff_calc* c) {
    if (!c) return fail(FDTD_ERR_ARG, "null handle");
    delete c;
    return FDTD_OK;
}

}  // extern "C"
