/* cpu_baseline.c -- CPU baseline engine timed beside the GPU.
 *
 * TEST/BENCH INFRASTRUCTURE ONLY (the cpu_baseline leg of bench.py and the
 * tests load it).  It restates the reference's fastest CPU engine so that the
 * GPU number can be quoted next to the reference algorithm run on the same
 * host:
 *   - storage of Engine_sse: 4-lane float vectors, z interleaved as
 *     z -> vector z % Nv, lane z / Nv (FDTD/engine_sse.h:36-46)
 *   - Engine_SSE_Compressed updates with the per-(x,y,zv) index into unique
 *     4-lane coefficient tuples and the SSE2 lane shifts at the z ends
 *     (FDTD/engine_sse_compressed.cpp:48-263, compression
 *     FDTD/operator_sse_compressed.cpp:120-176)
 *   - Engine_Multithread: x slabs per thread (FDTD/operator_multithread.cpp:89-106),
 *     a barrier after every extension call and after each main update,
 *     extensions without thread support run on thread 0
 *     (FDTD/engine_multithread.cpp:232-400, FDTD/extensions/engine_extension.cpp:50-90)
 *   - Engine_Ext_UPML split over threads along x (engine_ext_upml.cpp:46-55)
 *   - Engine_Ext_Excitation on thread 0 (engine_ext_excitation.cpp:33-94)
 *   - Engine_Ext_LorentzMaterial (Drude/Lorentz ADE) and Engine_Ext_TFSF on
 *     thread 0, as extensions without thread support run
 *     (engine_ext_lorentzmaterial.cpp:79-168, engine_ext_dispersive.cpp:76-127,
 *     engine_ext_tfsf.cpp:36-215)
 *   - FTZ|DAZ on every worker (tools/denormal.h:19-31)
 * Compiled -O3 -msse2 -ffp-contract=off.
 */
#include <emmintrin.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <xmmintrin.h>

typedef float v4sf __attribute__((vector_size(16)));

typedef struct {
    uint32_t start[3], n[3];
    float* c[6];          /* vv, vvfo, vvfn, ii, iifo, iifn [3][n] */
    float* vflux;         /* [3][n] */
    float* cflux;
    uint32_t tstart[256], tnum[256];
} upml_t;

typedef struct {
    size_t n;
    size_t* vec;          /* f4 index incl. component */
    int* lane;
    signed char* sign;
    int kind;
    double weight;
} probe_t;

/* one ADE order: cells as float offsets into V/I (component 0; + c*3 rows) */
typedef struct {
    uint32_t count, flags;   /* 1 volt, 2 curr, 4 volt Lorentz, 8 curr Lorentz */
    size_t* off;             /* [3][count] float offsets incl. component */
    float* k[6];             /* v_int, v_ext, v_lor, i_int, i_ext, i_lor [3][count] */
    float* s[4];             /* volt_ADE, volt_Lor, curr_ADE, curr_Lor [3][count] */
} ade_t;

/* TFSF injection list of one field */
typedef struct {
    uint32_t n, maxdelay;
    size_t* off;             /* float offsets incl. component */
    uint32_t* del;
    float *delta, *amp;
} tfsf_t;

typedef struct cpub {
    uint32_t N[3], Nv;
    size_t nrow;          /* Nx*Ny*Nv */
    v4sf *V, *I;
    uint32_t* idx;
    v4sf* tab[4][3];      /* vv, vi, iv, ii compressed per direction */
    uint32_t U;
    /* excitation */
    uint32_t exc_len, exc_period, nv;
    float *sig_v, *sig_i, *amp;
    uint32_t *pos, *dir, *del;
    /* Lorentz/Drude ADE orders, TFSF lists (volt, curr) */
    int n_ade;
    ade_t ade[8];
    tfsf_t tf[2];
    /* upml */
    int n_upml;
    upml_t upml[6];
    /* probes */
    int nprobes;
    probe_t probes[8];
    double* samples;      /* [nprobes][cap] */
    size_t nsamples, cap;
    uint32_t interval;
    /* threads */
    int nthreads;
    uint32_t tstart[256], tstop[256], tstop_h[256];
    pthread_barrier_t bar;
    uint32_t numTS;
    uint32_t iterTS;
} cpub;

static inline size_t row(const cpub* e, uint32_t n, uint32_t x, uint32_t y) {
    return (((size_t)n * e->N[0] + x) * e->N[1] + y) * e->Nv;
}
static inline float* elem(v4sf* F, const cpub* e, uint32_t n, uint32_t x, uint32_t y, uint32_t z) {
    return ((float*)&F[row(e, n, x, y) + z % e->Nv]) + z / e->Nv;
}

/* AssignJobs2Threads (tools/useful.cpp:45-75) */
static void assign_jobs(uint32_t jobs, uint32_t nthreads, uint32_t* jpt) {
    uint32_t per = jobs / nthreads;
    for (uint32_t n = 0; n < nthreads; ++n) { jpt[n] = per; jobs -= per; }
    for (uint32_t n = 0; n < nthreads; ++n)
        if (jobs > 0) { ++jpt[n]; --jobs; }
}

typedef struct { cpub* e; uint32_t x0, x1; } zjob_t;

static void* zero_worker(void* arg) {
    zjob_t* j = (zjob_t*)arg;
    cpub* e = j->e;
    const size_t a = (size_t)j->x0 * e->N[1] * e->Nv, b = (size_t)j->x1 * e->N[1] * e->Nv;
    for (int n = 0; n < 3; ++n) {
        memset(e->V + n * e->nrow + a, 0, (b - a) * 16);
        memset(e->I + n * e->nrow + a, 0, (b - a) * 16);
    }
    return NULL;
}

cpub* cpub_create(const uint32_t N[3], int nthreads) {
    cpub* e = (cpub*)calloc(1, sizeof(cpub));
    memcpy(e->N, N, 12);
    e->Nv = (N[2] + 3) / 4;
    e->nrow = (size_t)N[0] * N[1] * e->Nv;
    e->V = (v4sf*)aligned_alloc(64, (3 * e->nrow * 16 + 63) / 64 * 64);
    e->I = (v4sf*)aligned_alloc(64, (3 * e->nrow * 16 + 63) / 64 * 64);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    /* Operator_Multithread::CalcStartStopLines (FDTD/operator_multithread.cpp:89-106) */
    uint32_t jpt[256];
    uint32_t nt = (uint32_t)nthreads;
    assign_jobs(N[0], nt, jpt);
    while (nt > 1 && jpt[nt - 1] == 0) --nt;
    e->nthreads = (int)nt;
    e->tstart[0] = 0;
    e->tstop[0] = jpt[0] - 1;
    for (uint32_t n = 1; n < nt; ++n) { e->tstart[n] = e->tstop[n - 1] + 1; e->tstop[n] = e->tstart[n] + jpt[n] - 1; }
    for (uint32_t n = 0; n < nt; ++n) e->tstop_h[n] = (n == nt - 1) ? e->tstop[n] - 1 : e->tstop[n];
    pthread_barrier_init(&e->bar, NULL, nt);
    /* zero the fields from the threads that update them (first touch) */
    zjob_t zj[256];
    pthread_t th[256];
    for (uint32_t t = 0; t < nt; ++t) {
        zj[t].e = e; zj[t].x0 = e->tstart[t]; zj[t].x1 = e->tstop[t] + 1;
        pthread_create(&th[t], NULL, zero_worker, &zj[t]);
    }
    for (uint32_t t = 0; t < nt; ++t) pthread_join(th[t], NULL);
    return e;
}

void cpub_destroy(cpub* e) {
    if (!e) return;
    free(e->V); free(e->I); free(e->idx);
    for (int k = 0; k < 4; ++k) for (int n = 0; n < 3; ++n) free(e->tab[k][n]);
    free(e->sig_v); free(e->sig_i); free(e->amp); free(e->pos); free(e->dir); free(e->del);
    for (int a = 0; a < e->n_ade; ++a) {
        free(e->ade[a].off);
        for (int k = 0; k < 6; ++k) free(e->ade[a].k[k]);
        for (int k = 0; k < 4; ++k) free(e->ade[a].s[k]);
    }
    for (int w = 0; w < 2; ++w) { free(e->tf[w].off); free(e->tf[w].del); free(e->tf[w].delta); free(e->tf[w].amp); }
    for (int s = 0; s < e->n_upml; ++s) {
        for (int k = 0; k < 6; ++k) free(e->upml[s].c[k]);
        free(e->upml[s].vflux); free(e->upml[s].cflux);
    }
    for (int p = 0; p < e->nprobes; ++p) { free(e->probes[p].vec); free(e->probes[p].lane); free(e->probes[p].sign); }
    free(e->samples);
    pthread_barrier_destroy(&e->bar);
    free(e);
}

/* CompressOperator over 4-lane tuples of {vv,vi,iv,ii} x 3 (192 bytes),
 * first-occurrence numbering; equality is bytewise (memcmp) as in SSE_coeff */
typedef struct { float v[48]; } tuple_t;

static uint64_t hash_tuple(const tuple_t* t) {
    const uint32_t* w = (const uint32_t*)t->v;
    uint64_t h = 1469598103934665603ull;
    for (int i = 0; i < 48; ++i) { h ^= w[i]; h *= 1099511628211ull; }
    return h ^ (h >> 31);
}

int cpub_set_coeffs(cpub* e, const float* vv, const float* vi, const float* ii, const float* iv) {
    const size_t nd = (size_t)e->N[0] * e->N[1] * e->N[2];
    const float* src[4] = {vv, vi, iv, ii};   /* table order of the compressed operator */
    size_t cap = 1024;
    while (cap < 2 * e->nrow + 16) cap *= 2;
    int64_t* slots = (int64_t*)malloc(cap * sizeof(int64_t));
    for (size_t i = 0; i < cap; ++i) slots[i] = -1;
    tuple_t* uniq = (tuple_t*)malloc(sizeof(tuple_t) * 1024);
    size_t ucap = 1024, nu = 0;
    e->idx = (uint32_t*)malloc(e->nrow * sizeof(uint32_t));
    tuple_t t;
    for (uint32_t x = 0; x < e->N[0]; ++x)
        for (uint32_t y = 0; y < e->N[1]; ++y)
            for (uint32_t zv = 0; zv < e->Nv; ++zv) {
                memset(&t, 0, sizeof(t));
                for (int k = 0; k < 4; ++k)
                    for (int n = 0; n < 3; ++n)
                        for (int l = 0; l < 4; ++l) {
                            uint32_t z = zv + l * e->Nv;
                            if (z < e->N[2]) t.v[(k * 3 + n) * 4 + l] = src[k][n * nd + ((size_t)x * e->N[1] + y) * e->N[2] + z];
                        }
                uint64_t h = hash_tuple(&t);
                size_t s = h & (cap - 1);
                int64_t found = -1;
                while (slots[s] >= 0) {
                    if (memcmp(&uniq[slots[s]], &t, sizeof(t)) == 0) { found = slots[s]; break; }
                    s = (s + 1) & (cap - 1);
                }
                if (found < 0) {
                    if (nu == ucap) { ucap *= 2; uniq = (tuple_t*)realloc(uniq, ucap * sizeof(tuple_t)); }
                    uniq[nu] = t;
                    slots[s] = (int64_t)nu;
                    found = (int64_t)nu++;
                }
                e->idx[((size_t)x * e->N[1] + y) * e->Nv + zv] = (uint32_t)found;
            }
    free(slots);
    e->U = (uint32_t)nu;
    for (int k = 0; k < 4; ++k)
        for (int n = 0; n < 3; ++n) {
            e->tab[k][n] = (v4sf*)aligned_alloc(16, nu * 16);
            for (size_t u = 0; u < nu; ++u) memcpy(&e->tab[k][n][u], &uniq[u].v[(k * 3 + n) * 4], 16);
        }
    free(uniq);
    return (int)nu;
}

/* The same 4-lane compression from a per-cell compressed operator (index
 * [Nx][Ny][Nz] into rows of 12 values {vv0..2, vi0..2, ii0..2, iv0..2}, as
 * the streamed build gives it).  Per-cell rows are bytewise distinct, so two
 * 4-lane tuples are equal exactly when their four cell indices are; padding
 * lanes (z >= Nz) hold zeros and get index n_entries.  The 4-tuples are
 * deduplicated per thread over x slabs and merged; the numbering follows
 * the merge, which changes no arithmetic. */
typedef struct { uint32_t c[4]; } lanes4_t;

static uint64_t hash_quad(const lanes4_t* q) {
    uint64_t h = 1469598103934665603ull;
    for (int i = 0; i < 4; ++i) { h ^= q->c[i]; h *= 1099511628211ull; }
    return h ^ (h >> 29);
}

typedef struct {
    lanes4_t* keys; uint32_t* vals; size_t cap, n;   /* open addressing, vals = local id */
    lanes4_t* list;  size_t lcap;                    /* local id -> quad */
} qset_t;

static void qset_init(qset_t* q, size_t cap) {
    q->cap = cap; q->n = 0;
    q->keys = (lanes4_t*)malloc(cap * sizeof(lanes4_t));
    q->vals = (uint32_t*)malloc(cap * sizeof(uint32_t));
    for (size_t i = 0; i < cap; ++i) q->vals[i] = UINT32_MAX;
    q->lcap = 1024;
    q->list = (lanes4_t*)malloc(q->lcap * sizeof(lanes4_t));
}

static void qset_free(qset_t* q) { free(q->keys); free(q->vals); free(q->list); }

static uint32_t qset_get(qset_t* q, const lanes4_t* k) {
    if (2 * (q->n + 1) > q->cap) {   /* grow */
        qset_t g;
        qset_init(&g, 2 * q->cap);
        free(g.list);
        g.list = q->list; g.lcap = q->lcap; g.n = q->n;
        for (size_t i = 0; i < q->cap; ++i)
            if (q->vals[i] != UINT32_MAX) {
                size_t s = hash_quad(&q->keys[i]) & (g.cap - 1);
                while (g.vals[s] != UINT32_MAX) s = (s + 1) & (g.cap - 1);
                g.keys[s] = q->keys[i]; g.vals[s] = q->vals[i];
            }
        free(q->keys); free(q->vals);
        *q = g;
    }
    size_t s = hash_quad(k) & (q->cap - 1);
    while (q->vals[s] != UINT32_MAX) {
        if (memcmp(&q->keys[s], k, sizeof(lanes4_t)) == 0) return q->vals[s];
        s = (s + 1) & (q->cap - 1);
    }
    if (q->n == q->lcap) { q->lcap *= 2; q->list = (lanes4_t*)realloc(q->list, q->lcap * sizeof(lanes4_t)); }
    q->keys[s] = *k; q->vals[s] = (uint32_t)q->n; q->list[q->n] = *k;
    return (uint32_t)q->n++;
}

typedef struct { cpub* e; const uint32_t* cell; uint32_t pad_id, x0, x1; qset_t set; } qjob_t;

static void* quad_worker(void* arg) {
    qjob_t* j = (qjob_t*)arg;
    cpub* e = j->e;
    qset_init(&j->set, 4096);
    lanes4_t q;
    for (uint32_t x = j->x0; x < j->x1; ++x)
        for (uint32_t y = 0; y < e->N[1]; ++y) {
            const uint32_t* c = j->cell + ((size_t)x * e->N[1] + y) * e->N[2];
            for (uint32_t zv = 0; zv < e->Nv; ++zv) {
                for (int l = 0; l < 4; ++l) {
                    const uint32_t z = zv + l * e->Nv;
                    q.c[l] = z < e->N[2] ? c[z] : j->pad_id;
                }
                e->idx[((size_t)x * e->N[1] + y) * e->Nv + zv] = qset_get(&j->set, &q);
            }
        }
    return NULL;
}

typedef struct { cpub* e; uint32_t x0, x1; const uint32_t* remap; } rjob_t;

static void* remap_worker(void* arg) {
    rjob_t* j = (rjob_t*)arg;
    cpub* e = j->e;
    for (size_t p = (size_t)j->x0 * e->N[1] * e->Nv; p < (size_t)j->x1 * e->N[1] * e->Nv; ++p) e->idx[p] = j->remap[e->idx[p]];
    return NULL;
}

int cpub_set_coeffs_compressed(cpub* e, const uint32_t* cell, const double* table, uint32_t n_entries) {
    const int nt = e->nthreads;
    uint32_t jpt[256];
    assign_jobs(e->N[0], (uint32_t)nt, jpt);
    e->idx = (uint32_t*)malloc(e->nrow * sizeof(uint32_t));
    qjob_t* jobs = (qjob_t*)calloc((size_t)nt, sizeof(qjob_t));
    pthread_t th[256];
    uint32_t x = 0;
    for (int t = 0; t < nt; ++t) {
        jobs[t].e = e; jobs[t].cell = cell; jobs[t].pad_id = n_entries; jobs[t].x0 = x; jobs[t].x1 = x + jpt[t]; x += jpt[t];
        pthread_create(&th[t], NULL, quad_worker, &jobs[t]);
    }
    for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
    /* merge: global ids in thread order, then local -> global per thread */
    qset_t g;
    qset_init(&g, 4096);
    uint32_t* remaps[256];
    for (int t = 0; t < nt; ++t) {
        remaps[t] = (uint32_t*)malloc((jobs[t].set.n ? jobs[t].set.n : 1) * sizeof(uint32_t));
        for (size_t i = 0; i < jobs[t].set.n; ++i) remaps[t][i] = qset_get(&g, &jobs[t].set.list[i]);
    }
    rjob_t rj[256];
    for (int t = 0; t < nt; ++t) {
        rj[t].e = e; rj[t].x0 = jobs[t].x0; rj[t].x1 = jobs[t].x1; rj[t].remap = remaps[t];
        pthread_create(&th[t], NULL, remap_worker, &rj[t]);
    }
    for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
    const size_t nu = g.n;
    e->U = (uint32_t)nu;
    static const int col[4] = {0, 3, 9, 6};   /* tab order vv, vi, iv, ii -> table columns */
    for (int k = 0; k < 4; ++k)
        for (int n = 0; n < 3; ++n) {
            e->tab[k][n] = (v4sf*)aligned_alloc(16, (nu ? nu : 1) * 16);
            for (size_t u = 0; u < nu; ++u) {
                float* o = (float*)&e->tab[k][n][u];
                for (int l = 0; l < 4; ++l) {
                    const uint32_t c = g.list[u].c[l];
                    o[l] = c < n_entries ? (float)table[(size_t)c * 12 + col[k] + n] : 0.0f;
                }
            }
        }
    for (int t = 0; t < nt; ++t) { qset_free(&jobs[t].set); free(remaps[t]); }
    qset_free(&g);
    free(jobs);
    return (int)nu;
}

int cpub_set_excitation(cpub* e, uint32_t length, uint32_t period, const double* sig_v, uint32_t nv, const uint32_t* pos,
                        const uint32_t* dir, const double* amp, const uint32_t* del) {
    e->exc_len = length; e->exc_period = period; e->nv = nv;
    e->sig_v = (float*)malloc((length ? length : 1) * 4);
    for (uint32_t i = 0; i < length; ++i) e->sig_v[i] = (float)sig_v[i];
    e->amp = (float*)malloc((nv ? nv : 1) * 4);
    e->pos = (uint32_t*)malloc((nv ? 3 * (size_t)nv : 1) * 4);
    e->dir = (uint32_t*)malloc((nv ? nv : 1) * 4);
    e->del = (uint32_t*)malloc((nv ? nv : 1) * 4);
    for (uint32_t i = 0; i < nv; ++i) { e->amp[i] = (float)amp[i]; e->dir[i] = dir[i]; e->del[i] = del[i]; }
    memcpy(e->pos, pos, 3 * (size_t)nv * 4);
    return 0;
}

/* the current signal (the voltage TFSF hook reads it, engine_ext_tfsf.cpp:43-60) */
int cpub_set_signal_curr(cpub* e, uint32_t length, const double* sig_i) {
    free(e->sig_i);
    e->sig_i = (float*)malloc((length ? length : 1) * 4);
    for (uint32_t i = 0; i < length; ++i) e->sig_i[i] = (float)sig_i[i];
    return 0;
}

static size_t float_off(const cpub* e, uint32_t n, uint32_t x, uint32_t y, uint32_t z) {
    return (row(e, n, x, y) + z % e->Nv) * 4 + z / e->Nv;
}

/* one ADE order (Operator_Ext_LorentzMaterial tables): pos [3][count],
 * k lists [3][count] each (NULL = zeros), flags as above */
int cpub_add_ade(cpub* e, uint32_t count, const uint32_t* pos, const float* k0, const float* k1, const float* k2,
                 const float* k3, const float* k4, const float* k5, uint32_t flags) {
    if (e->n_ade >= 8) return -1;
    ade_t* a = &e->ade[e->n_ade++];
    const size_t c3 = 3 * (size_t)count;
    a->count = count; a->flags = flags;
    a->off = (size_t*)malloc((c3 ? c3 : 1) * sizeof(size_t));
    for (uint32_t i = 0; i < count; ++i)
        for (uint32_t c = 0; c < 3; ++c) a->off[c * (size_t)count + i] = float_off(e, c, pos[i], pos[count + i], pos[2 * (size_t)count + i]);
    const float* src[6] = {k0, k1, k2, k3, k4, k5};
    for (int k = 0; k < 6; ++k) {
        a->k[k] = (float*)calloc(c3 ? c3 : 1, 4);
        if (src[k]) memcpy(a->k[k], src[k], c3 * 4);
    }
    for (int k = 0; k < 4; ++k) a->s[k] = (float*)calloc(c3 ? c3 : 1, 4);
    return 0;
}

int cpub_set_tfsf(cpub* e, int which, uint32_t count, const uint32_t* pos, const uint32_t* comp, const uint32_t* delay,
                  const double* delta, const double* amp, uint32_t max_delay) {
    if (which < 0 || which > 1) return -1;
    tfsf_t* t = &e->tf[which];
    t->n = count; t->maxdelay = max_delay;
    t->off = (size_t*)malloc((count ? count : 1) * sizeof(size_t));
    t->del = (uint32_t*)malloc((count ? count : 1) * 4);
    t->delta = (float*)malloc((count ? count : 1) * 4);
    t->amp = (float*)malloc((count ? count : 1) * 4);
    for (uint32_t i = 0; i < count; ++i) {
        t->off[i] = float_off(e, comp[i], pos[i], pos[count + i], pos[2 * (size_t)count + i]);
        t->del[i] = delay[i];
        t->delta[i] = (float)delta[i];
        t->amp[i] = (float)amp[i];
    }
    return 0;
}

int cpub_add_upml(cpub* e, const uint32_t start[3], const uint32_t n[3], const float* vv, const float* vvfo, const float* vvfn,
                  const float* ii, const float* iifo, const float* iifn) {
    if (e->n_upml >= 6) return -1;
    upml_t* u = &e->upml[e->n_upml++];
    memcpy(u->start, start, 12);
    memcpy(u->n, n, 12);
    const size_t cnt = 3 * (size_t)n[0] * n[1] * n[2];
    const float* src[6] = {vv, vvfo, vvfn, ii, iifo, iifn};
    for (int k = 0; k < 6; ++k) { u->c[k] = (float*)malloc(cnt * 4); memcpy(u->c[k], src[k], cnt * 4); }
    u->vflux = (float*)calloc(cnt, 4);
    u->cflux = (float*)calloc(cnt, 4);
    /* Engine_Ext_UPML::SetNumberOfThreads: AssignJobs2Threads over the slab's x lines */
    uint32_t jpt[256];
    assign_jobs(n[0], (uint32_t)e->nthreads, jpt);
    uint32_t s = 0;
    for (int t = 0; t < e->nthreads; ++t) { u->tstart[t] = s; u->tnum[t] = jpt[t]; s += jpt[t]; }
    return 0;
}

/* probe edge lists (same construction as the device engine and the oracle) */
int cpub_add_probe(cpub* e, int kind, uint32_t nedges, const uint32_t* comp, const uint32_t* pos, const int* sign, double weight) {
    if (e->nprobes >= 8) return -1;
    probe_t* p = &e->probes[e->nprobes++];
    p->n = nedges; p->kind = kind; p->weight = weight;
    p->vec = (size_t*)malloc((nedges ? nedges : 1) * sizeof(size_t));
    p->lane = (int*)malloc((nedges ? nedges : 1) * sizeof(int));
    p->sign = (signed char*)malloc(nedges ? nedges : 1);
    for (uint32_t i = 0; i < nedges; ++i) {
        uint32_t x = pos[i], y = pos[nedges + i], z = pos[2 * nedges + i];
        p->vec[i] = row(e, comp[i], x, y) + z % e->Nv;
        p->lane[i] = (int)(z / e->Nv);
        p->sign[i] = (signed char)sign[i];
    }
    return e->nprobes - 1;
}

static void update_voltages(cpub* e, uint32_t startX, uint32_t numX) {
    const uint32_t Ny = e->N[1], Nv = e->Nv;
    v4sf *V0 = e->V, *V1 = e->V + e->nrow, *V2 = e->V + 2 * e->nrow;
    const v4sf *I0 = e->I, *I1 = e->I + e->nrow, *I2 = e->I + 2 * e->nrow;
    const v4sf *vv0 = e->tab[0][0], *vv1 = e->tab[0][1], *vv2 = e->tab[0][2];
    const v4sf *vi0 = e->tab[1][0], *vi1 = e->tab[1][1], *vi2 = e->tab[1][2];
    for (uint32_t x = startX; x < startX + numX; ++x) {
        const size_t sx = x > 0 ? (size_t)Ny * Nv : 0;
        for (uint32_t y = 0; y < Ny; ++y) {
            const size_t sy = y > 0 ? Nv : 0;
            const size_t r = ((size_t)x * Ny + y) * Nv;
            for (uint32_t zv = 1; zv < Nv; ++zv) {
                const size_t p = r + zv;
                const uint32_t c = e->idx[p];
                V0[p] *= vv0[c];
                V0[p] += vi0[c] * (I2[p] - I2[p - sy] - I1[p] + I1[p - 1]);
                V1[p] *= vv1[c];
                V1[p] += vi1[c] * (I0[p] - I0[p - 1] - I2[p] + I2[p - sx]);
                V2[p] *= vv2[c];
                V2[p] += vi2[c] * (I1[p] - I1[p - sx] - I0[p] + I0[p - sy]);
            }
            /* zv = 0: lane shift with zero fill (_mm_slli_si128 by 4 bytes) */
            const size_t p = r;
            const uint32_t c = e->idx[p];
            v4sf t1 = (v4sf)_mm_slli_si128((__m128i)I1[r + Nv - 1], 4);
            v4sf t0 = (v4sf)_mm_slli_si128((__m128i)I0[r + Nv - 1], 4);
            V0[p] *= vv0[c];
            V0[p] += vi0[c] * (I2[p] - I2[p - sy] - I1[p] + t1);
            V1[p] *= vv1[c];
            V1[p] += vi1[c] * (I0[p] - t0 - I2[p] + I2[p - sx]);
            V2[p] *= vv2[c];
            V2[p] += vi2[c] * (I1[p] - I1[p - sx] - I0[p] + I0[p - sy]);
        }
    }
}

static void update_currents(cpub* e, uint32_t startX, uint32_t numX) {
    const uint32_t Ny = e->N[1], Nv = e->Nv;
    const size_t sx = (size_t)Ny * Nv;
    const v4sf *V0 = e->V, *V1 = e->V + e->nrow, *V2 = e->V + 2 * e->nrow;
    v4sf *I0 = e->I, *I1 = e->I + e->nrow, *I2 = e->I + 2 * e->nrow;
    const v4sf *iv0 = e->tab[2][0], *iv1 = e->tab[2][1], *iv2 = e->tab[2][2];
    const v4sf *ii0 = e->tab[3][0], *ii1 = e->tab[3][1], *ii2 = e->tab[3][2];
    for (uint32_t x = startX; x < startX + numX; ++x)
        for (uint32_t y = 0; y + 1 < Ny; ++y) {
            const size_t r = ((size_t)x * Ny + y) * Nv;
            for (uint32_t zv = 0; zv + 1 < Nv; ++zv) {
                const size_t p = r + zv;
                const uint32_t c = e->idx[p];
                I0[p] *= ii0[c];
                I0[p] += iv0[c] * (V2[p] - V2[p + Nv] - V1[p] + V1[p + 1]);
                I1[p] *= ii1[c];
                I1[p] += iv1[c] * (V0[p] - V0[p + 1] - V2[p] + V2[p + sx]);
                I2[p] *= ii2[c];
                I2[p] += iv2[c] * (V1[p] - V1[p + sx] - V0[p] + V0[p + Nv]);
            }
            /* zv = Nv-1: lane shift right with zero fill (_mm_srli_si128) */
            const size_t p = r + Nv - 1;
            const uint32_t c = e->idx[p];
            v4sf t1 = (v4sf)_mm_srli_si128((__m128i)V1[r], 4);
            v4sf t0 = (v4sf)_mm_srli_si128((__m128i)V0[r], 4);
            I0[p] *= ii0[c];
            I0[p] += iv0[c] * (V2[p] - V2[p + Nv] - V1[p] + t1);
            I1[p] *= ii1[c];
            I1[p] += iv1[c] * (V0[p] - t0 - V2[p] + V2[p + sx]);
            I2[p] *= ii2[c];
            I2[p] += iv2[c] * (V1[p] - V1[p + sx] - V0[p] + V0[p + Nv]);
        }
}

static void upml_pre(cpub* e, upml_t* u, int tid, int curr) {
    const size_t cnt = (size_t)u->n[0] * u->n[1] * u->n[2];
    v4sf* F = curr ? e->I : e->V;
    const float* kv = u->c[curr ? 3 : 0];
    const float* kfo = u->c[curr ? 4 : 1];
    float* flux = curr ? u->cflux : u->vflux;
    for (uint32_t lx = u->tstart[tid]; lx < u->tstart[tid] + u->tnum[tid]; ++lx)
        for (uint32_t ly = 0; ly < u->n[1]; ++ly)
            for (uint32_t lz = 0; lz < u->n[2]; ++lz) {
                const size_t l = ((size_t)lx * u->n[1] + ly) * u->n[2] + lz;
                for (int c = 0; c < 3; ++c) {
                    float* f = elem(F, e, c, lx + u->start[0], ly + u->start[1], lz + u->start[2]);
                    const size_t q = c * cnt + l;
                    float h = kv[q] * *f - kfo[q] * flux[q];
                    *f = flux[q];
                    flux[q] = h;
                }
            }
}

static void upml_post(cpub* e, upml_t* u, int tid, int curr) {
    const size_t cnt = (size_t)u->n[0] * u->n[1] * u->n[2];
    v4sf* F = curr ? e->I : e->V;
    const float* kfn = u->c[curr ? 5 : 2];
    float* flux = curr ? u->cflux : u->vflux;
    for (uint32_t lx = u->tstart[tid]; lx < u->tstart[tid] + u->tnum[tid]; ++lx)
        for (uint32_t ly = 0; ly < u->n[1]; ++ly)
            for (uint32_t lz = 0; lz < u->n[2]; ++lz) {
                const size_t l = ((size_t)lx * u->n[1] + ly) * u->n[2] + lz;
                for (int c = 0; c < 3; ++c) {
                    float* f = elem(F, e, c, lx + u->start[0], ly + u->start[1], lz + u->start[2]);
                    const size_t q = c * cnt + l;
                    float h = flux[q];
                    flux[q] = *f;
                    *f = h + kfn[q] * flux[q];
                }
            }
}

static void excite(cpub* e) {
    int p = (int)e->numTS + 1;
    if (e->exc_period > 0) p = (int)e->exc_period;
    for (uint32_t i = 0; i < e->nv; ++i) {
        int k = (int)e->numTS - (int)e->del[i];
        k *= (k > 0);
        k %= p;
        k *= (k < (int)e->exc_len);
        float* f = elem(e->V, e, e->dir[i], e->pos[i], e->pos[e->nv + i], e->pos[2 * (size_t)e->nv + i]);
        *f = *f + e->amp[i] * e->sig_v[k];
    }
}

/* Engine_Ext_LorentzMaterial::DoPre{Voltage,Current}UpdatesImpl (engine_ext_lorentzmaterial.cpp:79-168) */
static void ade_pre(cpub* e, int curr) {
    float* F = (float*)(curr ? e->I : e->V);
    for (int o = 0; o < e->n_ade; ++o) {
        ade_t* a = &e->ade[o];
        if (!(a->flags & (curr ? 2u : 1u))) continue;
        const int lor = (a->flags & (curr ? 8u : 4u)) != 0;
        const float *kint = a->k[curr ? 3 : 0], *kext = a->k[curr ? 4 : 1], *klor = a->k[curr ? 5 : 2];
        float* J = a->s[curr ? 2 : 0];
        float* JL = a->s[curr ? 3 : 1];
        for (size_t i = 0; i < a->count; ++i)
            for (int c = 0; c < 3; ++c) {
                const size_t q = c * (size_t)a->count + i;
                const float f = F[a->off[q]];
                if (lor) {
                    JL[q] += klor[q] * J[q];
                    J[q] *= kint[q];
                    J[q] += kext[q] * (f - JL[q]);
                } else {
                    J[q] *= kint[q];
                    J[q] += kext[q] * f;
                }
            }
    }
}

/* Engine_Ext_Dispersive::Apply2{Voltages,Current}Impl (engine_ext_dispersive.cpp:76-127) */
static void ade_apply(cpub* e, int curr) {
    float* F = (float*)(curr ? e->I : e->V);
    for (int o = 0; o < e->n_ade; ++o) {
        ade_t* a = &e->ade[o];
        if (!(a->flags & (curr ? 2u : 1u))) continue;
        const float* J = a->s[curr ? 2 : 0];
        for (size_t i = 0; i < a->count; ++i)
            for (int c = 0; c < 3; ++c) {
                const size_t q = c * (size_t)a->count + i;
                F[a->off[q]] = F[a->off[q]] - J[q];
            }
    }
}

/* Engine_Ext_TFSF::DoPost{Voltage,Current}Updates (engine_ext_tfsf.cpp:36-215):
 * the voltage hook fills the delay lookup for n <= maxDelay, the current
 * hook for n < maxDelay, and reads the other field's signal */
static void tfsf(cpub* e, int curr) {
    const tfsf_t* t = &e->tf[curr];
    if (!t->n) return;
    static uint32_t L[65536];
    const uint32_t fill = curr ? t->maxdelay : t->maxdelay + 1;
    const uint32_t numTS = e->numTS, len = e->exc_len;
    const int per = (int)e->exc_period;
    for (uint32_t n = 0; n < fill && n < 65536; ++n) {
        if (numTS < n) L[n] = 0;
        else if ((numTS - n >= len) && (per == 0)) L[n] = 0;
        else L[n] = numTS - n;
        if (per > 0) L[n] = L[n] % (uint32_t)per;
    }
    const float* sig = curr ? e->sig_v : e->sig_i;
    float* F = (float*)(curr ? e->I : e->V);
    for (uint32_t i = 0; i < t->n; ++i) {
        const uint32_t d = t->del[i];
        const float dd = t->delta[i], a = t->amp[i];
        F[t->off[i]] = (float)((double)F[t->off[i]] + (1.0 - dd) * a * sig[L[d]] + dd * a * sig[L[d + 1]]);
    }
}

typedef struct { cpub* e; int tid; } targ_t;

/* NS_Engine_Multithread::thread::operator() loop body (FDTD/engine_multithread.cpp:308-400)
 * with the extension calls of :232-291, a barrier after each.  The extension
 * list sorted by priority is [UPML x n_upml, TFSF, Lorentz, Excitation]
 * (TFSF and Lorentz only when present); pre hooks run it in reverse. */
static void* worker(void* arg) {
    targ_t* a = (targ_t*)arg;
    cpub* e = a->e;
    const int tid = a->tid;
    _mm_setcsr(_mm_getcsr() | 0x8040);
    const uint32_t start = e->tstart[tid], stop = e->tstop[tid], stop_h = e->tstop_h[tid];
    const int has_tf = e->tf[0].n || e->tf[1].n, has_ade = e->n_ade > 0;
    for (uint32_t it = 0; it < e->iterTS; ++it) {
        /* DoPreVoltageUpdates: Excitation (no-op), Lorentz, TFSF (no-op), UPMLs */
        pthread_barrier_wait(&e->bar);
        if (has_ade) { if (tid == 0) ade_pre(e, 0); pthread_barrier_wait(&e->bar); }
        if (has_tf) pthread_barrier_wait(&e->bar);
        for (int s = e->n_upml - 1; s >= 0; --s) { upml_pre(e, &e->upml[s], tid, 0); pthread_barrier_wait(&e->bar); }
        update_voltages(e, start, stop - start + 1);
        pthread_barrier_wait(&e->bar);
        /* DoPostVoltageUpdates: UPMLs, TFSF, Lorentz (no-op), Excitation (no-op) */
        for (int s = 0; s < e->n_upml; ++s) { upml_post(e, &e->upml[s], tid, 0); pthread_barrier_wait(&e->bar); }
        if (has_tf) { if (tid == 0) tfsf(e, 0); pthread_barrier_wait(&e->bar); }
        if (has_ade) pthread_barrier_wait(&e->bar);
        pthread_barrier_wait(&e->bar);
        /* Apply2Voltages: UPMLs (no-op), TFSF (no-op), Lorentz, Excitation */
        for (int s = 0; s < e->n_upml; ++s) pthread_barrier_wait(&e->bar);
        if (has_tf) pthread_barrier_wait(&e->bar);
        if (has_ade) { if (tid == 0) ade_apply(e, 0); pthread_barrier_wait(&e->bar); }
        if (tid == 0) excite(e);
        pthread_barrier_wait(&e->bar);
        /* current half step, same structure */
        pthread_barrier_wait(&e->bar);
        if (has_ade) { if (tid == 0) ade_pre(e, 1); pthread_barrier_wait(&e->bar); }
        if (has_tf) pthread_barrier_wait(&e->bar);
        for (int s = e->n_upml - 1; s >= 0; --s) { upml_pre(e, &e->upml[s], tid, 1); pthread_barrier_wait(&e->bar); }
        if (stop_h + 1 > start) update_currents(e, start, stop_h - start + 1);
        pthread_barrier_wait(&e->bar);
        for (int s = 0; s < e->n_upml; ++s) { upml_post(e, &e->upml[s], tid, 1); pthread_barrier_wait(&e->bar); }
        if (has_tf) { if (tid == 0) tfsf(e, 1); pthread_barrier_wait(&e->bar); }
        if (has_ade) pthread_barrier_wait(&e->bar);
        pthread_barrier_wait(&e->bar);
        for (int s = 0; s < e->n_upml; ++s) pthread_barrier_wait(&e->bar);
        if (has_tf) pthread_barrier_wait(&e->bar);
        if (has_ade) { if (tid == 0) ade_apply(e, 1); pthread_barrier_wait(&e->bar); }
        pthread_barrier_wait(&e->bar);                 /* Excitation Apply2Current (no current sources) */
        if (tid == 0) ++e->numTS;
    }
    return NULL;
}

static void sample(cpub* e) {
    if (!e->interval || e->numTS % e->interval) return;
    if (e->nsamples == e->cap) {
        e->cap = e->cap ? 2 * e->cap : 1024;
        e->samples = (double*)realloc(e->samples, e->cap * 8 * sizeof(double));
    }
    for (int k = 0; k < e->nprobes; ++k) {
        probe_t* p = &e->probes[k];
        double val = 0;
        if (p->kind == 0) {
            for (size_t i = 0; i < p->n; ++i) {
                double f = ((float*)&e->V[p->vec[i]])[p->lane[i]];
                if (p->sign[i] > 0) val += f; else val -= f;
            }
        } else {
            float acc = 0;
            for (size_t i = 0; i < p->n; ++i) {
                float f = ((float*)&e->I[p->vec[i]])[p->lane[i]];
                if (p->sign[i] > 0) acc += f; else acc -= f;
            }
            val = acc;
        }
        e->samples[e->nsamples * 8 + k] = val * p->weight;
    }
    e->nsamples++;
}

/* Runs nsteps as RunFDTD does: IterateTS(interval) then Process() (probes on
 * the main thread); returns the wall time of the loop in seconds. */
double cpub_run(cpub* e, uint32_t nsteps, uint32_t interval) {
    e->interval = interval;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    if (e->numTS == 0) sample(e);
    uint32_t left = nsteps;
    while (left) {
        uint32_t chunk = interval ? interval - e->numTS % interval : left;
        if (chunk > left) chunk = left;
        e->iterTS = chunk;
        pthread_t th[256];
        targ_t args[256];
        for (int t = 1; t < e->nthreads; ++t) { args[t].e = e; args[t].tid = t; pthread_create(&th[t], NULL, worker, &args[t]); }
        args[0].e = e; args[0].tid = 0;
        worker(&args[0]);
        for (int t = 1; t < e->nthreads; ++t) pthread_join(th[t], NULL);
        left -= chunk;
        sample(e);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

int cpub_get_field(cpub* e, int which, double* out) {
    v4sf* F = which ? e->I : e->V;
    for (uint32_t n = 0; n < 3; ++n)
        for (uint32_t x = 0; x < e->N[0]; ++x)
            for (uint32_t y = 0; y < e->N[1]; ++y)
                for (uint32_t z = 0; z < e->N[2]; ++z)
                    *out++ = (double)*elem(F, e, n, x, y, z);
    return 0;
}

uint32_t cpub_num_unique(cpub* e) { return e->U; }
int cpub_num_threads(cpub* e) { return e->nthreads; }
uint32_t cpub_num_timesteps(cpub* e) { return e->numTS; }

int cpub_read_samples(cpub* e, int probe, double* out, uint32_t max_n) {
    uint32_t n = e->nsamples < max_n ? (uint32_t)e->nsamples : max_n;
    for (uint32_t i = 0; i < n; ++i) out[i] = e->samples[i * 8 + probe];
    return (int)e->nsamples;
}
