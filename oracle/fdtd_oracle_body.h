/* fdtd_oracle_body.h -- engine body of the CPU oracle, included twice by
 * fdtd_oracle.c with REAL = float (suffix _f) and REAL = double (suffix _d).
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker the device engine is compared
 * against; nothing in the product links or calls it.
 *
 * Each function restates the reference at the cited lines; loops run in the
 * reference's order and every floating-point operation is written in the
 * reference's order so the float build reproduces the reference's bits.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SUF)

typedef struct {
    REAL *vv, *vi, *ii, *iv;          /* [3][N] */
    REAL *V, *I;                      /* fields [3][N] */
    /* UPML slabs */
    uint32_t n_upml;
    uint32_t ustart[6][3], un[6][3];
    REAL* ucoef[6][6];                /* slab, array (vv,vvfo,vvfn,ii,iifo,iifn) [3][n] */
    REAL* vflux[6];
    REAL* cflux[6];
    /* excitation */
    uint32_t exc_len, exc_period;
    REAL *sig_v, *sig_i;
    uint32_t nv, nc;
    uint32_t *v_pos, *v_dir, *v_del, *c_pos, *c_dir, *c_del;
    REAL *v_amp, *c_amp;
    /* ADE orders */
    uint32_t n_ade;
    uint32_t ade_count[8], ade_flags[8];
    uint32_t* ade_pos[8];             /* [3][count] */
    REAL* ade_k[8][6];                /* v_int, v_ext, v_lor, i_int, i_ext, i_lor [3][count] */
    REAL* ade_s[8][4];                /* volt_ADE, volt_Lor, curr_ADE, curr_Lor [3][count] */
    /* TFSF */
    uint32_t tf_n[2], tf_maxdelay[2];
    uint32_t *tf_pos[2], *tf_comp[2], *tf_del[2];
    REAL *tf_delta[2], *tf_amp[2];
    /* Mur ABC faces, ascending face (2*axis + top) */
    uint32_t n_mur;
    int mur_face[6];
    uint32_t mur_start[6];
    REAL* mur_k[6];                   /* [2][na][nb]: nyP, nyPP */
    REAL* mur_buf[6];
    /* lumped RLC cells */
    uint32_t n_rlc;
    uint32_t *rlc_pos, *rlc_dir;      /* [3][n], [n] */
    REAL* rlc_c[9];                   /* ilv, i2v, vv2, vj1, vj2, vvd, ib0, b1, b2 */
    REAL* rlc_il;
    REAL* rlc_vd[3];                  /* rotated once per step, as the reference rotates its pointers */
    REAL* rlc_j[3];
    /* steady-state detection */
    uint32_t ss_n, ss_p;
    uint32_t *ss_pos, *ss_dir;        /* [3][n], [n] */
    double* ss_rec;                   /* [n][2p] */
    double ss_last_diff, ss_last_energy;
} FN(fields_t);

/* Releases every array the fields own (free(NULL) is a no-op). */
static void FN(fields_free)(FN(fields_t) * f) {
    free(f->vv); free(f->vi); free(f->ii); free(f->iv); free(f->V); free(f->I);
    for (int s = 0; s < 6; ++s) {
        for (int k = 0; k < 6; ++k) free(f->ucoef[s][k]);
        free(f->vflux[s]); free(f->cflux[s]);
        free(f->mur_k[s]); free(f->mur_buf[s]);
    }
    free(f->sig_v); free(f->sig_i);
    free(f->v_pos); free(f->v_dir); free(f->v_del); free(f->c_pos); free(f->c_dir); free(f->c_del);
    free(f->v_amp); free(f->c_amp);
    for (int a = 0; a < 8; ++a) {
        free(f->ade_pos[a]);
        for (int k = 0; k < 6; ++k) free(f->ade_k[a][k]);
        for (int k = 0; k < 4; ++k) free(f->ade_s[a][k]);
    }
    for (int w = 0; w < 2; ++w) {
        free(f->tf_pos[w]); free(f->tf_comp[w]); free(f->tf_del[w]); free(f->tf_delta[w]); free(f->tf_amp[w]);
    }
    free(f->rlc_pos); free(f->rlc_dir);
    for (int k = 0; k < 9; ++k) free(f->rlc_c[k]);
    free(f->rlc_il);
    for (int k = 0; k < 3; ++k) { free(f->rlc_vd[k]); free(f->rlc_j[k]); }
    free(f->ss_pos); free(f->ss_dir); free(f->ss_rec);
}

#define IDX(x, y, z) (((size_t)(x) * N1 + (y)) * N2 + (z))

/* Engine::UpdateVoltages (FDTD/engine.cpp:108-161).  The z-1 neighbour at
 * z = 0 is the zero-filled lane of Engine_SSE_Compressed::UpdateVoltages
 * (FDTD/engine_sse_compressed.cpp:95-153); x-1 and y-1 at index 0 use the
 * cell itself (shift[] is a bool). */
static void FN(update_voltages)(FN(fields_t) * f, const uint32_t* Nl) {
    const size_t N0 = Nl[0], N1 = Nl[1], N2 = Nl[2], nd = N0 * N1 * N2;
    REAL *V0 = f->V, *V1 = f->V + nd, *V2 = f->V + 2 * nd;
    const REAL *I0 = f->I, *I1 = f->I + nd, *I2 = f->I + 2 * nd;
    for (size_t x = 0; x < N0; ++x) {
        const size_t sx = x > 0 ? 1 : 0;
        for (size_t y = 0; y < N1; ++y) {
            const size_t sy = y > 0 ? 1 : 0;
            for (size_t z = 0; z < N2; ++z) {
                const size_t p = IDX(x, y, z);
                const REAL i1_zm = z > 0 ? I1[p - 1] : (REAL)0;
                const REAL i0_zm = z > 0 ? I0[p - 1] : (REAL)0;
                V0[p] *= f->vv[p];
                V0[p] += f->vi[p] * (I2[p] - I2[IDX(x, y - sy, z)] - I1[p] + i1_zm);
                V1[p] *= f->vv[nd + p];
                V1[p] += f->vi[nd + p] * (I0[p] - i0_zm - I2[p] + I2[IDX(x - sx, y, z)]);
                V2[p] *= f->vv[2 * nd + p];
                V2[p] += f->vi[2 * nd + p] * (I1[p] - I1[IDX(x - sx, y, z)] - I0[p] + I0[IDX(x, y - sy, z)]);
            }
        }
    }
}

/* Engine::UpdateCurrents (FDTD/engine.cpp:163-210) over x<Nx-1, y<Ny-1, z<Nz-1 */
static void FN(update_currents)(FN(fields_t) * f, const uint32_t* Nl) {
    const size_t N0 = Nl[0], N1 = Nl[1], N2 = Nl[2], nd = N0 * N1 * N2;
    const REAL *V0 = f->V, *V1 = f->V + nd, *V2 = f->V + 2 * nd;
    REAL *I0 = f->I, *I1 = f->I + nd, *I2 = f->I + 2 * nd;
    for (size_t x = 0; x + 1 < N0; ++x)
        for (size_t y = 0; y + 1 < N1; ++y)
            for (size_t z = 0; z + 1 < N2; ++z) {
                const size_t p = IDX(x, y, z);
                I0[p] *= f->ii[p];
                I0[p] += f->iv[p] * (V2[p] - V2[IDX(x, y + 1, z)] - V1[p] + V1[p + 1]);
                I1[p] *= f->ii[nd + p];
                I1[p] += f->iv[nd + p] * (V0[p] - V0[p + 1] - V2[p] + V2[IDX(x + 1, y, z)]);
                I2[p] *= f->ii[2 * nd + p];
                I2[p] += f->iv[2 * nd + p] * (V1[p] - V1[IDX(x + 1, y, z)] - V0[p] + V0[IDX(x, y + 1, z)]);
            }
}

/* Engine_Ext_UPML::DoPre{Voltage,Current}UpdatesImpl (engine_ext_upml.cpp:57-98,141-182) */
static void FN(upml_pre)(FN(fields_t) * f, const uint32_t* Nl, int curr) {
    const size_t N1 = Nl[1], N2 = Nl[2], nd = (size_t)Nl[0] * N1 * N2;
    for (uint32_t s = 0; s < f->n_upml; ++s) {
        const uint32_t *st = f->ustart[s], *n = f->un[s];
        const size_t cnt = (size_t)n[0] * n[1] * n[2];
        const REAL* kv = f->ucoef[s][curr ? 3 : 0];
        const REAL* kfo = f->ucoef[s][curr ? 4 : 1];
        REAL* flux = curr ? f->cflux[s] : f->vflux[s];
        REAL* F = curr ? f->I : f->V;
        for (uint32_t lx = 0; lx < n[0]; ++lx)
            for (uint32_t ly = 0; ly < n[1]; ++ly)
                for (uint32_t lz = 0; lz < n[2]; ++lz) {
                    const size_t l = ((size_t)lx * n[1] + ly) * n[2] + lz;
                    const size_t p = IDX(lx + st[0], ly + st[1], lz + st[2]);
                    for (int c = 0; c < 3; ++c) {
                        const size_t q = c * cnt + l;
                        REAL h = kv[q] * F[c * nd + p] - kfo[q] * flux[q];
                        F[c * nd + p] = flux[q];
                        flux[q] = h;
                    }
                }
    }
}

/* Engine_Ext_UPML::DoPost{Voltage,Current}UpdatesImpl (engine_ext_upml.cpp:100-135,184-234) */
static void FN(upml_post)(FN(fields_t) * f, const uint32_t* Nl, int curr) {
    const size_t N1 = Nl[1], N2 = Nl[2], nd = (size_t)Nl[0] * N1 * N2;
    for (uint32_t s = 0; s < f->n_upml; ++s) {
        const uint32_t *st = f->ustart[s], *n = f->un[s];
        const size_t cnt = (size_t)n[0] * n[1] * n[2];
        const REAL* kfn = f->ucoef[s][curr ? 5 : 2];
        REAL* flux = curr ? f->cflux[s] : f->vflux[s];
        REAL* F = curr ? f->I : f->V;
        for (uint32_t lx = 0; lx < n[0]; ++lx)
            for (uint32_t ly = 0; ly < n[1]; ++ly)
                for (uint32_t lz = 0; lz < n[2]; ++lz) {
                    const size_t l = ((size_t)lx * n[1] + ly) * n[2] + lz;
                    const size_t p = IDX(lx + st[0], ly + st[1], lz + st[2]);
                    for (int c = 0; c < 3; ++c) {
                        const size_t q = c * cnt + l;
                        REAL h = flux[q];
                        flux[q] = F[c * nd + p];
                        F[c * nd + p] = h + kfn[q] * flux[q];
                    }
                }
    }
}

/* Engine_Ext_LorentzMaterial::DoPre{Voltage,Current}UpdatesImpl (engine_ext_lorentzmaterial.cpp:79-168) */
static void FN(ade_pre)(FN(fields_t) * f, const uint32_t* Nl, int curr) {
    const size_t N1 = Nl[1], N2 = Nl[2], nd = (size_t)Nl[0] * N1 * N2;
    const REAL* F = curr ? f->I : f->V;
    for (uint32_t o = 0; o < f->n_ade; ++o) {
        const uint32_t fl = f->ade_flags[o];
        if (!(fl & (curr ? 2u : 1u))) continue;
        const int lor = (fl & (curr ? 8u : 4u)) != 0;
        const size_t cnt = f->ade_count[o];
        const uint32_t* pos = f->ade_pos[o];
        const REAL* kint = f->ade_k[o][curr ? 3 : 0];
        const REAL* kext = f->ade_k[o][curr ? 4 : 1];
        const REAL* klor = f->ade_k[o][curr ? 5 : 2];
        REAL* J = f->ade_s[o][curr ? 2 : 0];
        REAL* JL = f->ade_s[o][curr ? 3 : 1];
        for (size_t i = 0; i < cnt; ++i) {
            const size_t p = IDX(pos[i], pos[cnt + i], pos[2 * cnt + i]);
            for (int c = 0; c < 3; ++c) {
                const size_t q = c * cnt + i;
                if (lor) {
                    JL[q] += klor[q] * J[q];
                    J[q] *= kint[q];
                    J[q] += kext[q] * (F[c * nd + p] - JL[q]);
                } else {
                    J[q] *= kint[q];
                    J[q] += kext[q] * F[c * nd + p];
                }
            }
        }
    }
}

/* Engine_Ext_Dispersive::Apply2{Voltages,Current}Impl (engine_ext_dispersive.cpp:76-127) */
static void FN(ade_apply)(FN(fields_t) * f, const uint32_t* Nl, int curr) {
    const size_t N1 = Nl[1], N2 = Nl[2], nd = (size_t)Nl[0] * N1 * N2;
    REAL* F = curr ? f->I : f->V;
    for (uint32_t o = 0; o < f->n_ade; ++o) {
        if (!(f->ade_flags[o] & (curr ? 2u : 1u))) continue;
        const size_t cnt = f->ade_count[o];
        const uint32_t* pos = f->ade_pos[o];
        const REAL* J = f->ade_s[o][curr ? 2 : 0];
        for (size_t i = 0; i < cnt; ++i) {
            const size_t p = IDX(pos[i], pos[cnt + i], pos[2 * cnt + i]);
            for (int c = 0; c < 3; ++c) F[c * nd + p] = F[c * nd + p] - J[c * cnt + i];
        }
    }
}

/* Engine_Ext_Excitation::Apply2{Voltages,Current}Impl (engine_ext_excitation.cpp:33-94) */
static void FN(excite)(FN(fields_t) * f, const uint32_t* Nl, int curr, uint32_t numTS) {
    const size_t N1 = Nl[1], N2 = Nl[2], nd = (size_t)Nl[0] * N1 * N2;
    const uint32_t cnt = curr ? f->nc : f->nv;
    const uint32_t *pos = curr ? f->c_pos : f->v_pos, *dir = curr ? f->c_dir : f->v_dir, *del = curr ? f->c_del : f->v_del;
    const REAL* amp = curr ? f->c_amp : f->v_amp;
    const REAL* sig = curr ? f->sig_i : f->sig_v;
    REAL* F = curr ? f->I : f->V;
    int p = (int)numTS + 1;
    if (f->exc_period > 0) p = (int)f->exc_period;
    for (uint32_t i = 0; i < cnt; ++i) {
        int k = (int)numTS - (int)del[i];
        k *= (k > 0);
        k %= p;
        k *= (k < (int)f->exc_len);
        const size_t q = dir[i] * nd + IDX(pos[i], pos[cnt + i], pos[2 * (size_t)cnt + i]);
        F[q] = F[q] + amp[i] * sig[k];
    }
}

/* Engine_Ext_TFSF::DoPost{Voltage,Current}Updates (engine_ext_tfsf.cpp:36-215) */
static void FN(tfsf)(FN(fields_t) * f, const uint32_t* Nl, int curr, uint32_t numTS) {
    const size_t N1 = Nl[1], N2 = Nl[2], nd = (size_t)Nl[0] * N1 * N2;
    const uint32_t cnt = f->tf_n[curr];
    if (!cnt) return;
    const uint32_t maxd = f->tf_maxdelay[curr];
    /* the voltage hook fills n <= maxDelay, the current hook n < maxDelay (:43,133) */
    const uint32_t fill = curr ? maxd : maxd + 1;
    static uint32_t lookup_store[2][65536];
    uint32_t* L = lookup_store[sizeof(REAL) == 8];
    const uint32_t len = f->exc_len;
    const int per = (int)f->exc_period;
    for (uint32_t n = 0; n < fill && n < 65536; ++n) {
        if (numTS < n) L[n] = 0;
        else if ((numTS - n >= len) && (per == 0)) L[n] = 0;
        else L[n] = numTS - n;
        if (per > 0) L[n] = L[n] % (uint32_t)per;
    }
    const REAL* sig = curr ? f->sig_v : f->sig_i;
    REAL* F = curr ? f->I : f->V;
    for (uint32_t i = 0; i < cnt; ++i) {
        const uint32_t* pos = f->tf_pos[curr];
        const size_t q = f->tf_comp[curr][i] * nd + IDX(pos[i], pos[cnt + i], pos[2 * (size_t)cnt + i]);
        const uint32_t d = f->tf_del[curr][i];
        const REAL dd = f->tf_delta[curr][i], a = f->tf_amp[curr][i];
        F[q] = (REAL)((double)F[q] + (1.0 - dd) * a * sig[L[d]] + dd * a * sig[L[d + 1]]);
    }
}

/* Engine_Ext_Mur_ABC (engine_ext_mur_abc.cpp:83-168): phase 0 DoPreVoltageUpdates,
 * 1 DoPostVoltageUpdates, 2 Apply2Voltages; each face acts from its start
 * step on (IsActive, engine_ext_mur_abc.h:55).  Faces run in ascending order
 * for the pre hook and in descending order for post and apply: the
 * engine's extension list holds equal-priority extensions in reverse
 * creation order (FDTD/engine.cpp:84-88) and openems.cpp:386-397 creates
 * them x-lo .. z-hi. */
static void FN(mur)(FN(fields_t) * f, const uint32_t* Nl, uint32_t numTS, int phase) {
    const size_t N1 = Nl[1], N2 = Nl[2], nd = (size_t)Nl[0] * N1 * N2;
    for (uint32_t q = 0; q < f->n_mur; ++q) {
        const uint32_t s = phase == 0 ? q : f->n_mur - 1 - q;
        if (numTS < f->mur_start[s]) continue;
        const int face = f->mur_face[s], ny = face / 2, nyP = (ny + 1) % 3, nyPP = (ny + 2) % 3;
        const uint32_t line = face % 2 ? Nl[ny] - 1 : 0, shift = face % 2 ? Nl[ny] - 2 : 1;
        const size_t na = Nl[nyP], nb = Nl[nyPP];
        uint32_t pos[3], pss[3];
        pos[ny] = line;
        pss[ny] = shift;
        for (uint32_t a = 0; a < na; ++a) {
            pos[nyP] = pss[nyP] = a;
            for (uint32_t b = 0; b < nb; ++b) {
                pos[nyPP] = pss[nyPP] = b;
                const size_t p = IDX(pos[0], pos[1], pos[2]), ps = IDX(pss[0], pss[1], pss[2]);
                for (int t = 0; t < 2; ++t) {
                    REAL* V = f->V + (size_t)(t ? nyPP : nyP) * nd;
                    const size_t j = (size_t)t * na * nb + (size_t)a * nb + b;
                    const REAL k = f->mur_k[s][j];
                    if (phase == 0) f->mur_buf[s][j] = V[ps] - k * V[p];
                    else if (phase == 1) f->mur_buf[s][j] += k * V[ps];
                    else V[p] = f->mur_buf[s][j];
                }
            }
        }
    }
}

/* Engine_Ext_LumpedRLC::DoPreVoltageUpdates / Apply2VoltagesImpl
 * (engine_ext_lumpedRLC.cpp:81-149) */
static void FN(rlc_pre)(FN(fields_t) * f) {
    if (!f->n_rlc) return;
    REAL* t = f->rlc_vd[2];
    f->rlc_vd[2] = f->rlc_vd[1];
    f->rlc_vd[1] = f->rlc_vd[0];
    f->rlc_vd[0] = t;
    for (uint32_t i = 0; i < f->n_rlc; ++i) f->rlc_il[i] += (f->rlc_c[1][i]) * (f->rlc_c[0][i]) * f->rlc_vd[1][i];
}

static void FN(rlc_apply)(FN(fields_t) * f, const uint32_t* Nl) {
    if (!f->n_rlc) return;
    const size_t N1 = Nl[1], N2 = Nl[2], nd = (size_t)Nl[0] * N1 * N2, n = f->n_rlc;
    REAL* t = f->rlc_j[2];
    f->rlc_j[2] = f->rlc_j[1];
    f->rlc_j[1] = f->rlc_j[0];
    f->rlc_j[0] = t;
    REAL **vd = f->rlc_vd, **j = f->rlc_j, **c = f->rlc_c;
    const uint32_t* p = f->rlc_pos;
    for (size_t i = 0; i < n; ++i)
        vd[0][i] = f->V[f->rlc_dir[i] * nd + IDX(p[i], p[n + i], p[2 * n + i])];
    for (size_t i = 0; i < n; ++i) {
        vd[0][i] = (c[5][i]) * (vd[0][i] - f->rlc_il[i] + (c[2][i]) * vd[2][i] + (c[3][i]) * j[1][i] + (c[4][i]) * j[2][i]);
        j[0][i] = (c[6][i]) * (vd[0][i] - vd[2][i]) - ((c[7][i]) * (c[6][i])) * j[1][i] - ((c[8][i]) * (c[6][i])) * j[2][i];
    }
    for (size_t i = 0; i < n; ++i)
        f->V[f->rlc_dir[i] * nd + IDX(p[i], p[n + i], p[2 * n + i])] = vd[0][i];
}

/* Engine_Interface_SSE_FDTD::CalcFastEnergy (FDTD/engine_interface_sse_fdtd.cpp:37-69):
 * eps0*sum(V^2) + mu0*sum(I^2) over x < Nx-1, y < Ny-1 and every z, here
 * accumulated in double in x, y, z, component order (the reference sums in
 * four float lanes; the engine under test accumulates in FP64 too). */
static double FN(fast_energy)(const FN(fields_t) * f, const uint32_t* Nl) {
    const size_t N0 = Nl[0], N1 = Nl[1], N2 = Nl[2], nd = N0 * N1 * N2;
    double e = 0, h = 0;
    for (size_t x = 0; x + 1 < N0; ++x)
        for (size_t y = 0; y + 1 < N1; ++y)
            for (size_t z = 0; z < N2; ++z)
                for (int c = 0; c < 3; ++c) {
                    const size_t q = c * nd + IDX(x, y, z);
                    const double v = (double)f->V[q], i = (double)f->I[q];
                    e += v * v;
                    h += i * i;
                }
    return 8.85418781762e-12 * e + 1.256637062e-6 * h;
}

/* Engine_Ext_SteadyState::Apply2Voltages (engine_ext_steadystate.cpp:51-117):
 * record every probe at slot numTS % 2p; at numTS % p == 0, numTS >= 2p,
 * compare the energy with the last evaluation and, per probe, the window at
 * new_pos with the one at old_pos (new_pos = 0 whenever numTS % 2p <= p,
 * which both possible values satisfy). */
static void FN(steady_state)(FN(fields_t) * f, const uint32_t* Nl, uint32_t numTS) {
    if (!f->ss_p) return;
    const size_t N1 = Nl[1], N2 = Nl[2], nd = (size_t)Nl[0] * N1 * N2;
    const uint32_t p = f->ss_p, n = f->ss_n, rel = numTS % (2 * p);
    const uint32_t* q = f->ss_pos;
    for (uint32_t i = 0; i < n; ++i)
        f->ss_rec[(size_t)i * 2 * p + rel] = (double)f->V[f->ss_dir[i] * nd + IDX(q[i], q[n + i], q[2 * n + i])];
    if (numTS % p != 0 || numTS < 2 * p) return;
    int no_valid = 1;
    double diff = 0;
    const double energy = FN(fast_energy)(f, Nl);
    if (f->ss_last_energy > 0) {
        diff = fabs(energy - f->ss_last_energy) / f->ss_last_energy;
        no_valid = 0;
    }
    f->ss_last_energy = energy;
    uint32_t old_pos = 0, new_pos = p;
    if (rel <= p) { new_pos = 0; old_pos = p; }
    double* cp = (double*)malloc((n ? n : 1) * sizeof(double));
    double* dp = (double*)malloc((n ? n : 1) * sizeof(double));
    double max_pow = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const double* b = f->ss_rec + (size_t)i * 2 * p;
        cp[i] = 0;
        dp[i] = 0;
        for (uint32_t t = 0; t < p; ++t) {
            cp[i] += b[t + new_pos] * b[t + new_pos];
            dp[i] += (b[t + old_pos] - b[t + new_pos]) * (b[t + old_pos] - b[t + new_pos]);
        }
        if (max_pow < cp[i]) max_pow = cp[i];
    }
    for (uint32_t i = 0; i < n; ++i)
        if (cp[i] > max_pow * 1e-2) {
            if (diff < dp[i] / cp[i]) diff = dp[i] / cp[i];
            no_valid = 0;
        }
    if (no_valid || diff > 1) diff = 1;
    f->ss_last_diff = diff;
    free(cp);
    free(dp);
}

/* One timestep in the order of Engine::IterateTS (FDTD/engine.cpp:255-274) with
 * the extensions sorted by priority: pre hooks Excitation, Mur, Lorentz, RLC,
 * TFSF, UPML; post and apply hooks steady state, UPML, TFSF, RLC, Lorentz,
 * Mur, Excitation (the steady-state extension has only an apply hook). */
static void FN(step)(FN(fields_t) * f, const uint32_t* Nl, uint32_t numTS) {
    FN(mur)(f, Nl, numTS, 0);
    FN(ade_pre)(f, Nl, 0);
    FN(rlc_pre)(f);
    FN(upml_pre)(f, Nl, 0);
    FN(update_voltages)(f, Nl);
    FN(upml_post)(f, Nl, 0);
    FN(tfsf)(f, Nl, 0, numTS);
    FN(mur)(f, Nl, numTS, 1);
    FN(steady_state)(f, Nl, numTS);
    FN(rlc_apply)(f, Nl);
    FN(ade_apply)(f, Nl, 0);
    FN(mur)(f, Nl, numTS, 2);
    FN(excite)(f, Nl, 0, numTS);

    FN(ade_pre)(f, Nl, 1);
    FN(upml_pre)(f, Nl, 1);
    FN(update_currents)(f, Nl);
    FN(upml_post)(f, Nl, 1);
    FN(tfsf)(f, Nl, 1, numTS);
    FN(ade_apply)(f, Nl, 1);
    FN(excite)(f, Nl, 1, numTS);
}

/* ProcessCurrent::CalcIntegral accumulates in FDTD_FLOAT (Common/processcurrent.cpp:95) */
static double FN(current_sum)(const REAL* I, size_t n, const size_t* off, const signed char* sg) {
    REAL acc = 0;
    for (size_t e = 0; e < n; ++e) {
        if (sg[e] > 0) acc += I[off[e]];
        else acc -= I[off[e]];
    }
    return (double)acc;
}

#undef IDX
#undef FN
#undef CAT
#undef CAT2
