/* fdtd_oracle.c -- CPU oracle for the device FDTD engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This restates the reference engine
 * (openEMS Engine / Engine_SSE_Compressed with its extensions and probes) in
 * plain C so the HIP engine can be checked against it; only tests/, the smoke
 * check and the CPU-baseline leg of bench.py load it, and never as the thing
 * measured or shipped.
 *
 * Build: see oracle/Makefile (-O2 -ffp-contract=off, no -ffast-math: every
 * float operation rounds as in the reference, which is compiled for baseline
 * x86-64 without FMA).  With ftz=1 the float engine runs with MXCSR FTZ|DAZ
 * like the SSE engines (tools/denormal.h:19-31, FDTD/engine_sse.cpp:40).
 *
 * Pinning (what checks each part, see tests/test_oracle.py):
 *   update_voltages/currents  closed-form PEC cavity resonances, energy
 *                             conservation of a lossless box, brute-force
 *                             pure-Python step on tiny grids
 *   upml_pre/post             reduces to the plain update when kappa = 0
 *                             (disabled cells: vv' = 0, vvfo = 0, vvfn = 1),
 *                             normal-incidence reflection below -40 dB
 *   excite                    index clamp rules of engine_ext_excitation.cpp:50-58
 *                             against hand-computed cases
 *   ade_pre/apply             single-cell ODE of the Drude circuit in closed form
 *   voltage/current probes    hand-summed edge lists on fixed fields
 *   energy                    exact sum on small grids
 *   tfsf                      scattered-field leakage of plane-wave boxes and the
 *                             closed-form incident wave (tests/test_tfsf.py)
 *   mur                       plane-wave reflection of a parallel-plate line against
 *                             a PEC end and a 3x longer line (tests/test_mur.py)
 *   rlc_pre/apply             discrete parallel-LC frequency, series-loop frequency
 *                             and decay (tests/test_rlc.py)
 *   steady_state              closed-form windows and energy ratios of a linear
 *                             ramp, an exactly periodic field, a sine-driven line
 *                             that settles (tests/test_steady_state.py)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <xmmintrin.h>

#define REAL float
#define SUF _f
#include "fdtd_oracle_body.h"
#undef REAL
#undef SUF
#define REAL double
#define SUF _d
#include "fdtd_oracle_body.h"
#undef REAL
#undef SUF

typedef struct {
    int kind;
    int norm_dir;
    double weight;
    uint32_t td_interval, fd_interval;
    uint32_t n_freq;
    double* freqs;
    size_t n_edges;
    size_t* off;      /* element offsets incl. component */
    signed char* sign;
    double* td;       /* (t,val) pairs */
    size_t td_n, td_cap;
    double* fd;       /* (re,im) per freq */
} probe_t;

typedef struct orc {
    uint32_t N[3];
    double dT;
    int dbl;
    int ftz;
    uint32_t numTS;
    int sampled0;
    fields_t_f f32;
    fields_t_d f64;
    probe_t probes[64];
    int nprobes;
} orc;

static size_t nd_of(const orc* o) { return (size_t)o->N[0] * o->N[1] * o->N[2]; }

static void* copy_as(const orc* o, int dtype, const void* src, size_t n) {
    if (!src) return NULL;
    if (o->dbl) {
        double* d = (double*)malloc(n * sizeof(double));
        for (size_t i = 0; i < n; ++i) d[i] = dtype ? ((const double*)src)[i] : (double)((const float*)src)[i];
        return d;
    }
    float* d = (float*)malloc(n * sizeof(float));
    for (size_t i = 0; i < n; ++i) d[i] = dtype ? (float)((const double*)src)[i] : ((const float*)src)[i];
    return d;
}

static void* zeros(const orc* o, size_t n) { return calloc(n, o->dbl ? sizeof(double) : sizeof(float)); }

orc* orc_create(const uint32_t N[3], double dT, int use_double, int ftz) {
    orc* o = (orc*)calloc(1, sizeof(orc));
    memcpy(o->N, N, sizeof(o->N));
    o->dT = dT;
    o->dbl = use_double != 0;
    o->ftz = ftz != 0;
    const size_t n3 = 3 * nd_of(o);
    if (o->dbl) { o->f64.V = (double*)zeros(o, n3); o->f64.I = (double*)zeros(o, n3); }
    else { o->f32.V = (float*)zeros(o, n3); o->f32.I = (float*)zeros(o, n3); }
    return o;
}

void orc_destroy(orc* o) {
    if (!o) return;
    fields_free_f(&o->f32);
    fields_free_d(&o->f64);
    for (int p = 0; p < o->nprobes; ++p) {
        free(o->probes[p].freqs); free(o->probes[p].off); free(o->probes[p].sign);
        free(o->probes[p].td); free(o->probes[p].fd);
    }
    free(o);
}

int orc_set_coeffs(orc* o, int dtype, const void* vv, const void* vi, const void* ii, const void* iv) {
    const size_t n3 = 3 * nd_of(o);
    if (o->dbl) {
        o->f64.vv = (double*)copy_as(o, dtype, vv, n3); o->f64.vi = (double*)copy_as(o, dtype, vi, n3);
        o->f64.ii = (double*)copy_as(o, dtype, ii, n3); o->f64.iv = (double*)copy_as(o, dtype, iv, n3);
    } else {
        o->f32.vv = (float*)copy_as(o, dtype, vv, n3); o->f32.vi = (float*)copy_as(o, dtype, vi, n3);
        o->f32.ii = (float*)copy_as(o, dtype, ii, n3); o->f32.iv = (float*)copy_as(o, dtype, iv, n3);
    }
    return 0;
}

static uint32_t* copy_u32(const uint32_t* s, size_t n) {
    uint32_t* d = (uint32_t*)malloc((n ? n : 1) * 4);
    if (n) memcpy(d, s, n * 4);
    return d;
}

int orc_set_excitation(orc* o, uint32_t length, uint32_t period, const double* sv, const double* si,
                       uint32_t nv, const uint32_t* vpos, const uint32_t* vdir, const double* vamp, const uint32_t* vdel,
                       uint32_t nc, const uint32_t* cpos, const uint32_t* cdir, const double* camp, const uint32_t* cdel) {
#define SETEXC(F, T)                                                     \
    F.exc_len = length; F.exc_period = period;                           \
    F.sig_v = (T*)copy_as(o, 1, sv, length); F.sig_i = (T*)copy_as(o, 1, si, length); \
    F.nv = nv; F.nc = nc;                                                \
    F.v_pos = copy_u32(vpos, 3 * (size_t)nv); F.v_dir = copy_u32(vdir, nv); F.v_del = copy_u32(vdel, nv); \
    F.v_amp = (T*)copy_as(o, 1, vamp, nv);                               \
    F.c_pos = copy_u32(cpos, 3 * (size_t)nc); F.c_dir = copy_u32(cdir, nc); F.c_del = copy_u32(cdel, nc); \
    F.c_amp = (T*)copy_as(o, 1, camp, nc);
    if (o->dbl) { SETEXC(o->f64, double) } else { SETEXC(o->f32, float) }
#undef SETEXC
    return 0;
}

int orc_add_upml(orc* o, const uint32_t start[3], const uint32_t n[3], int dtype, const void* vv, const void* vvfo,
                 const void* vvfn, const void* ii, const void* iifo, const void* iifn) {
    const void* src[6] = {vv, vvfo, vvfn, ii, iifo, iifn};
    const size_t cnt = 3 * (size_t)n[0] * n[1] * n[2];
#define ADDU(F, T)                                                           \
    {                                                                        \
        uint32_t s = F.n_upml;                                               \
        if (s >= 6) return -1;                                               \
        memcpy(F.ustart[s], start, 12); memcpy(F.un[s], n, 12);             \
        for (int k = 0; k < 6; ++k) F.ucoef[s][k] = (T*)copy_as(o, dtype, src[k], cnt); \
        F.vflux[s] = (T*)zeros(o, cnt); F.cflux[s] = (T*)zeros(o, cnt);     \
        F.n_upml = s + 1;                                                    \
    }
    if (o->dbl) ADDU(o->f64, double) else ADDU(o->f32, float)
#undef ADDU
    return 0;
}

int orc_add_ade(orc* o, uint32_t count, const uint32_t* pos, int dtype, const void* v_int, const void* v_ext, const void* v_lor,
                const void* i_int, const void* i_ext, const void* i_lor, uint32_t flags) {
    const void* src[6] = {v_int, v_ext, v_lor, i_int, i_ext, i_lor};
    const size_t c3 = 3 * (size_t)count;
#define ADDA(F, T)                                                             \
    {                                                                          \
        uint32_t a = F.n_ade;                                                  \
        if (a >= 8) return -1;                                                 \
        F.ade_count[a] = count; F.ade_flags[a] = flags;                        \
        F.ade_pos[a] = copy_u32(pos, c3);                                      \
        for (int k = 0; k < 6; ++k) F.ade_k[a][k] = src[k] ? (T*)copy_as(o, dtype, src[k], c3) : (T*)zeros(o, c3); \
        for (int k = 0; k < 4; ++k) F.ade_s[a][k] = (T*)zeros(o, c3);         \
        F.n_ade = a + 1;                                                       \
    }
    if (o->dbl) ADDA(o->f64, double) else ADDA(o->f32, float)
#undef ADDA
    return 0;
}

int orc_add_mur(orc* o, int face, int dtype, const void* k_nyP, const void* k_nyPP, uint32_t start_ts) {
    if (face < 0 || face > 5) return -1;
    const int ny = face / 2;
    const size_t plane = (size_t)o->N[(ny + 1) % 3] * o->N[(ny + 2) % 3];
#define ADDM(F, T)                                                                     \
    {                                                                                  \
        uint32_t s = F.n_mur;                                                          \
        if (s >= 6) return -1;                                                         \
        while (s > 0 && F.mur_face[s - 1] > face) {                                    \
            F.mur_face[s] = F.mur_face[s - 1]; F.mur_start[s] = F.mur_start[s - 1];     \
            F.mur_k[s] = F.mur_k[s - 1]; F.mur_buf[s] = F.mur_buf[s - 1];               \
            --s;                                                                       \
        }                                                                              \
        T* k = (T*)malloc(2 * plane * sizeof(T));                                      \
        T* a = (T*)copy_as(o, dtype, k_nyP, plane);                                    \
        T* b = (T*)copy_as(o, dtype, k_nyPP, plane);                                   \
        memcpy(k, a, plane * sizeof(T)); memcpy(k + plane, b, plane * sizeof(T));      \
        free(a); free(b);                                                              \
        F.mur_face[s] = face; F.mur_start[s] = start_ts; F.mur_k[s] = k;               \
        F.mur_buf[s] = (T*)zeros(o, 2 * plane);                                        \
        ++F.n_mur;                                                                     \
    }
    if (o->dbl) ADDM(o->f64, double) else ADDM(o->f32, float)
#undef ADDM
    return 0;
}

int orc_set_rlc(orc* o, uint32_t count, const uint32_t* pos, const uint32_t* dir, int dtype, const void* coef) {
#define SETR(F, T)                                                                     \
    {                                                                                  \
        F.n_rlc = count;                                                               \
        F.rlc_pos = copy_u32(pos, 3 * (size_t)count);                                  \
        F.rlc_dir = copy_u32(dir, count);                                              \
        T* all = (T*)copy_as(o, dtype, coef, 9 * (size_t)count);                       \
        for (int k = 0; k < 9; ++k) {                                                  \
            F.rlc_c[k] = (T*)malloc(count * sizeof(T));                                \
            memcpy(F.rlc_c[k], all + (size_t)k * count, count * sizeof(T));            \
        }                                                                              \
        free(all);                                                                     \
        F.rlc_il = (T*)zeros(o, count);                                                \
        for (int k = 0; k < 3; ++k) { F.rlc_vd[k] = (T*)zeros(o, count); F.rlc_j[k] = (T*)zeros(o, count); } \
    }
    if (!count) return 0;
    if (o->dbl) SETR(o->f64, double) else SETR(o->f32, float)
#undef SETR
    return 0;
}

int orc_set_steady_state(orc* o, uint32_t period_ts, uint32_t count, const uint32_t* pos, const uint32_t* dir) {
    if (!period_ts) return -1;
#define SETS(F)                                                                  \
    {                                                                            \
        free(F.ss_pos); free(F.ss_dir); free(F.ss_rec);                          \
        F.ss_n = count; F.ss_p = period_ts;                                      \
        F.ss_pos = copy_u32(pos, 3 * (size_t)count);                             \
        F.ss_dir = copy_u32(dir, count);                                         \
        F.ss_rec = (double*)calloc((count ? count : 1) * 2 * (size_t)period_ts, sizeof(double)); \
        F.ss_last_diff = 1; F.ss_last_energy = 0;                                \
    }
    if (o->dbl) SETS(o->f64) else SETS(o->f32)
#undef SETS
    return 0;
}

double orc_steady_state_last_diff(orc* o) { return o->dbl ? o->f64.ss_last_diff : o->f32.ss_last_diff; }

int orc_set_tfsf(orc* o, int which, uint32_t count, const uint32_t* pos, const uint32_t* comp, const uint32_t* delay,
                 const double* delta, const double* amp, uint32_t max_delay) {
#define SETT(F, T)                                                         \
    F.tf_n[which] = count; F.tf_maxdelay[which] = max_delay;               \
    F.tf_pos[which] = copy_u32(pos, 3 * (size_t)count);                    \
    F.tf_comp[which] = copy_u32(comp, count); F.tf_del[which] = copy_u32(delay, count); \
    F.tf_delta[which] = (T*)copy_as(o, 1, delta, count); F.tf_amp[which] = (T*)copy_as(o, 1, amp, count);
    if (o->dbl) { SETT(o->f64, double) } else { SETT(o->f32, float) }
#undef SETT
    return 0;
}

static void add_edge(orc* o, probe_t* p, int n, uint32_t x, uint32_t y, uint32_t z, int sgn) {
    p->off = (size_t*)realloc(p->off, (p->n_edges + 1) * sizeof(size_t));
    p->sign = (signed char*)realloc(p->sign, p->n_edges + 1);
    p->off[p->n_edges] = (size_t)n * nd_of(o) + ((size_t)x * o->N[1] + y) * o->N[2] + z;
    p->sign[p->n_edges] = (signed char)sgn;
    p->n_edges++;
}

/* Engine_Interface_FDTD::CalcVoltageIntegral (FDTD/engine_interface_fdtd.cpp:200-226) and
 * ProcessCurrent::CalcIntegral (Common/processcurrent.cpp:93-168) edge lists */
int orc_add_probe(orc* o, int kind, const uint32_t a[3], const uint32_t b[3], int norm_dir, uint32_t inside_mask, double weight,
                  uint32_t td_interval, uint32_t fd_interval, const double* freqs, uint32_t n_freq) {
    if (o->nprobes >= 64) return -1;
    probe_t* p = &o->probes[o->nprobes];
    memset(p, 0, sizeof(*p));
    p->kind = kind; p->norm_dir = norm_dir; p->weight = weight;
    p->td_interval = td_interval; p->fd_interval = fd_interval; p->n_freq = n_freq;
    if (n_freq) {
        p->freqs = (double*)malloc(n_freq * sizeof(double));
        memcpy(p->freqs, freqs, n_freq * sizeof(double));
        p->fd = (double*)calloc(2 * n_freq, sizeof(double));
    }
    if (kind == 0) {
        for (int n = 0; n < 3; ++n) {
            if (a[n] < b[n]) {
                uint32_t pos[3] = {a[0], a[1], a[2]};
                for (; pos[n] < b[n]; ++pos[n]) add_edge(o, p, n, pos[0], pos[1], pos[2], +1);
            } else if (a[n] > b[n]) {
                uint32_t pos[3] = {b[0], b[1], b[2]};
                for (; pos[n] < a[n]; ++pos[n]) add_edge(o, p, n, pos[0], pos[1], pos[2], -1);
            }
        }
    } else {
        const int si0 = inside_mask & 1, si1 = (inside_mask >> 1) & 1, si2 = (inside_mask >> 2) & 1;
        const int so0 = (inside_mask >> 3) & 1, so1 = (inside_mask >> 4) & 1, so2 = (inside_mask >> 5) & 1;
        uint32_t i;
        switch (norm_dir) {
        case 0:
            if (so0 && si2) for (i = a[1] + 1; i <= b[1]; ++i) add_edge(o, p, 1, b[0], i, a[2], +1);
            if (so0 && so1) for (i = a[2] + 1; i <= b[2]; ++i) add_edge(o, p, 2, b[0], b[1], i, +1);
            if (si0 && so2) for (i = a[1] + 1; i <= b[1]; ++i) add_edge(o, p, 1, a[0], i, b[2], -1);
            if (si0 && si1) for (i = a[2] + 1; i <= b[2]; ++i) add_edge(o, p, 2, a[0], a[1], i, -1);
            break;
        case 1:
            if (si0 && si1) for (i = a[2] + 1; i <= b[2]; ++i) add_edge(o, p, 2, a[0], a[1], i, +1);
            if (so1 && so2) for (i = a[0] + 1; i <= b[0]; ++i) add_edge(o, p, 0, i, b[1], b[2], +1);
            if (so0 && so1) for (i = a[2] + 1; i <= b[2]; ++i) add_edge(o, p, 2, b[0], b[1], i, -1);
            if (si1 && si2) for (i = a[0] + 1; i <= b[0]; ++i) add_edge(o, p, 0, i, a[1], a[2], -1);
            break;
        default:
            if (si1 && si2) for (i = a[0] + 1; i <= b[0]; ++i) add_edge(o, p, 0, i, a[1], a[2], +1);
            if (so0 && si2) for (i = a[1] + 1; i <= b[1]; ++i) add_edge(o, p, 1, b[0], i, a[2], +1);
            if (so1 && so2) for (i = a[0] + 1; i <= b[0]; ++i) add_edge(o, p, 0, i, b[1], b[2], -1);
            if (si0 && so2) for (i = a[1] + 1; i <= b[1]; ++i) add_edge(o, p, 1, a[0], i, b[2], -1);
            break;
        }
    }
    return o->nprobes++;
}

/* ProcessIntegral::Process (Common/processintegral.cpp:134-171) at the current numTS */
static void sample(orc* o) {
    const uint32_t ts = o->numTS;
    for (int k = 0; k < o->nprobes; ++k) {
        probe_t* p = &o->probes[k];
        const int td = p->td_interval && ts % p->td_interval == 0;
        const int fd = p->fd_interval && ts % p->fd_interval == 0;
        if (!td && !fd) continue;
        double val = 0;
        if (p->kind == 0) {
            for (size_t e = 0; e < p->n_edges; ++e) {
                const double f = o->dbl ? o->f64.V[p->off[e]] : (double)o->f32.V[p->off[e]];
                if (p->sign[e] > 0) val += f; else val -= f;
            }
        } else {
            val = o->dbl ? current_sum_d(o->f64.I, p->n_edges, p->off, p->sign) : current_sum_f(o->f32.I, p->n_edges, p->off, p->sign);
        }
        const double t = ((double)ts + (p->kind == 1 ? 0.5 : 0.0)) * o->dT;
        if (td) {
            if (p->td_n == p->td_cap) {
                p->td_cap = p->td_cap ? 2 * p->td_cap : 256;
                p->td = (double*)realloc(p->td, p->td_cap * 2 * sizeof(double));
            }
            p->td[2 * p->td_n] = t;
            p->td[2 * p->td_n + 1] = val * p->weight;
            p->td_n++;
        }
        if (fd) {
            for (uint32_t q = 0; q < p->n_freq; ++q) {
                const double ang = -2.0 * M_PI * p->freqs[q] * t;
                const double g = val * p->weight;
                double re = g * cos(ang), im = g * sin(ang);
                re = re * 2.0 * o->dT * (double)p->fd_interval;
                im = im * 2.0 * o->dT * (double)p->fd_interval;
                p->fd[2 * q] += re;
                p->fd[2 * q + 1] += im;
            }
        }
    }
}

int orc_iterate(orc* o, uint32_t nsteps) {
    unsigned int old = _mm_getcsr();
    if (o->ftz && !o->dbl) _mm_setcsr(old | 0x8040);
    if (!o->sampled0) { sample(o); o->sampled0 = 1; }
    for (uint32_t s = 0; s < nsteps; ++s) {
        if (o->dbl) step_d(&o->f64, o->N, o->numTS);
        else step_f(&o->f32, o->N, o->numTS);
        ++o->numTS;
        sample(o);
    }
    _mm_setcsr(old);
    return 0;
}

uint32_t orc_num_timesteps(orc* o) { return o->numTS; }

int orc_get_field(orc* o, int which, double* out) {
    const size_t n3 = 3 * nd_of(o);
    if (o->dbl) memcpy(out, which ? o->f64.I : o->f64.V, n3 * sizeof(double));
    else {
        const float* s = which ? o->f32.I : o->f32.V;
        for (size_t i = 0; i < n3; ++i) out[i] = (double)s[i];
    }
    return 0;
}

int orc_set_field(orc* o, int which, const double* in) {
    const size_t n3 = 3 * nd_of(o);
    if (o->dbl) memcpy(which ? o->f64.I : o->f64.V, in, n3 * sizeof(double));
    else {
        float* d = which ? o->f32.I : o->f32.V;
        for (size_t i = 0; i < n3; ++i) d[i] = (float)in[i];
    }
    return 0;
}

int orc_read_probe(orc* o, int id, double* t, double* v, uint32_t max_n) {
    if (id < 0 || id >= o->nprobes) return -1;
    probe_t* p = &o->probes[id];
    uint32_t n = p->td_n < max_n ? (uint32_t)p->td_n : max_n;
    for (uint32_t i = 0; i < n; ++i) { t[i] = p->td[2 * i]; v[i] = p->td[2 * i + 1]; }
    return (int)p->td_n;
}

int orc_read_probe_fd(orc* o, int id, double* re, double* im) {
    if (id < 0 || id >= o->nprobes) return -1;
    probe_t* p = &o->probes[id];
    for (uint32_t q = 0; q < p->n_freq; ++q) { re[q] = p->fd[2 * q]; im[q] = p->fd[2 * q + 1]; }
    return 0;
}

/* eps0*sum(V^2) + mu0*sum(I^2) over x<Nx-1, y<Ny-1, all z, in double */
double orc_energy(orc* o) {
    return o->dbl ? fast_energy_d(&o->f64, o->N) : fast_energy_f(&o->f32, o->N);
}
