/*
 * wipems_fdtd.h -- C ABI of the MI355X (gfx950) FDTD engine.
 *
 * This is the boundary an openEMS-style host plugs a device engine into.  The
 * entry points replace the hot-path pieces of the reference:
 *
 *   fdtd_create / fdtd_destroy      Engine::New / Engine::Init / Engine::Reset
 *                                   (FDTD/engine.cpp:25-57,98-106), allocation
 *                                   of the volt/curr arrays
 *   fdtd_set_coeffs                 Operator vv/vi/ii/iv arrays
 *                                   (FDTD/operator.h:63-78, FDTD/operator_sse.h:76-80)
 *   fdtd_set_coeffs_compressed      Operator_SSE_Compressed::m_Op_index and
 *                                   f4_{vv,vi,iv,ii}_Compressed
 *                                   (FDTD/operator_sse_compressed.h:76-80)
 *   fdtd_set_excitation             Operator_Ext_Excitation arrays + Excitation signals
 *                                   (FDTD/extensions/operator_ext_excitation.h:68-81,
 *                                    FDTD/excitation.cpp:151-177), consumed by
 *                                   Engine_Ext_Excitation (engine_ext_excitation.cpp:33-94)
 *   fdtd_add_upml                   Operator_Ext_UPML vv/vvfo/vvfn/ii/iifo/iifn
 *                                   (FDTD/extensions/operator_ext_upml.h:98-103),
 *                                   consumed by Engine_Ext_UPML (engine_ext_upml.cpp:57-234)
 *   fdtd_add_ade                    Operator_Ext_LorentzMaterial per-order lists
 *                                   (FDTD/extensions/operator_ext_lorentzmaterial.h:49-59),
 *                                   consumed by Engine_Ext_LorentzMaterial / _Dispersive
 *                                   (engine_ext_lorentzmaterial.cpp:79-168, engine_ext_dispersive.cpp:76-127)
 *   fdtd_set_tfsf                   Operator_Ext_TFSF delay/amp tables
 *                                   (FDTD/extensions/operator_ext_tfsf.h:71-78),
 *                                   consumed by Engine_Ext_TFSF (engine_ext_tfsf.cpp:36-215)
 *   fdtd_iterate                    Engine::IterateTS (FDTD/engine.cpp:255-274)
 *   fdtd_get_num_timesteps          Engine::GetNumberOfTimesteps (FDTD/engine.h:48)
 *   fdtd_get_value / fdtd_set_value Engine::GetVolt/GetCurr/SetVolt/SetCurr (FDTD/engine.h:52-61)
 *   fdtd_get_field / fdtd_set_field bulk form of the same accessors (field dumps)
 *   fdtd_add_probe, fdtd_read_probe ProcessVoltage / ProcessCurrent integrals and
 *                                   ProcessIntegral::Process sampling + FD accumulation
 *                                   (Common/processintegral.cpp:134-171,
 *                                    FDTD/engine_interface_fdtd.cpp:200-226,
 *                                    Common/processcurrent.cpp:93-168)
 *   fdtd_calc_energy                Engine_Interface_SSE_FDTD::CalcFastEnergy
 *                                   (FDTD/engine_interface_sse_fdtd.cpp:37-69), in FP64
 *   fdtd_set_edge_lengths,          ProcessFieldsFD (Common/processfields_fd.cpp:72-107) over
 *   fdtd_add_dft_box, fdtd_read_dft a box of mesh nodes with NODE_INTERPOLATE
 *                                   (FDTD/engine_interface_fdtd.cpp:63-158), FP64 values
 *                                   and accumulators; six of them make the nf2ff box
 *                                   (python/openEMS/nf2ff.py:61-77)
 *   fdtd_add_td_box                 ProcessFieldsTD (Common/processfields_td.cpp:54-91)
 *   fdtd_add_dft_dump_box,          the same for the dump types E/H/J/rotH/D/B and the
 *   fdtd_add_td_dump_box,           three interpolations (Common/processfields.cpp:280-400,
 *   fdtd_set_material               FDTD/engine_interface_fdtd.cpp:60-288) with the
 *                                   operator's stored materials (FDTD/operator.cpp:1176-1188)
 *
 * Conventions: every function returns 0 on success and a negative FDTD_ERR_*
 * code on failure; fdtd_last_error() returns the message of the last failure
 * on the calling thread.  No C++ types and no exceptions cross this boundary.
 * The caller owns every host buffer it passes; the library owns device memory
 * and copies what it is given at call time.  One handle is used by one host
 * thread at a time; work is ordered on the handle's own HIP stream.
 *
 * Array layout at this boundary is the reference's basic layout:
 * field/coefficient arrays are [3][Nx][Ny][Nz] row-major (z fastest),
 * matching volt[n][x][y][z] of FDTD/engine.h:95-96.
 */
#ifndef WIPEMS_FDTD_H
#define WIPEMS_FDTD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define FDTD_OK            0
#define FDTD_ERR_ARG      -1   /* invalid argument                          */
#define FDTD_ERR_HIP      -2   /* HIP runtime error                         */
#define FDTD_ERR_STATE    -3   /* call not valid in the current state       */
#define FDTD_ERR_RANGE    -4   /* index or size out of range                */
#define FDTD_ERR_NOMEM    -5   /* device or host allocation failed          */
#define FDTD_ERR_NODEVICE -6   /* no usable gfx950 device                   */

/* ---- field storage precision of a handle ---- */
#define FDTD_PREC_FP32 0       /* float fields, float coefficients (the drop-in)      */
#define FDTD_PREC_FP64 1       /* double fields and coefficients (FP64 oracle engine)  */
#define FDTD_PREC_FP16 2       /* half fields with per-component scales, FP32 math     */
#define FDTD_PREC_BF16 3       /* bfloat16 fields with per-component scales, FP32 math */

/* ---- element type tag of host arrays passed in or out ---- */
#define FDTD_DTYPE_F32 0
#define FDTD_DTYPE_F64 1

/* ---- which field ---- */
#define FDTD_VOLT 0
#define FDTD_CURR 1

/* ---- probe kinds ---- */
#define FDTD_PROBE_VOLTAGE 0   /* ProcessVoltage: line integral of volt           */
#define FDTD_PROBE_CURRENT 1   /* ProcessCurrent: loop integral of curr, dual time */

typedef struct fdtd_engine fdtd_engine;

typedef struct fdtd_grid_desc {
    uint32_t num_lines[3];     /* mesh lines per axis (Operator::GetNumberOfLines) */
    double   dT;               /* timestep in seconds (Operator::GetTimestep)      */
    int32_t  precision;        /* FDTD_PREC_*                                       */
    uint32_t flags;            /* reserved, pass 0                                  */
} fdtd_grid_desc;

/* Excitation: signals and the sparse soft-source edge lists.  Arrays of
 * positions are given as x[], y[], z[] blocks of length count ([3][count]). */
typedef struct fdtd_excitation {
    uint32_t       length;      /* Excitation::GetLength                          */
    uint32_t       period;      /* signal period in timesteps, 0 = not periodic   */
    const double*  signal_volt; /* [length] Excitation::GetVoltageSignal          */
    const double*  signal_curr; /* [length] Excitation::GetCurrentSignal          */
    uint32_t       volt_count;
    const uint32_t* volt_pos;   /* [3][volt_count]                                */
    const uint32_t* volt_dir;   /* [volt_count]                                   */
    const double*  volt_amp;    /* [volt_count]                                   */
    const uint32_t* volt_delay; /* [volt_count] in timesteps                      */
    uint32_t       curr_count;
    const uint32_t* curr_pos;
    const uint32_t* curr_dir;
    const double*  curr_amp;
    const uint32_t* curr_delay;
} fdtd_excitation;

/* ---- lifetime ---- */
int  fdtd_create(const fdtd_grid_desc* desc, fdtd_engine** out);
int  fdtd_destroy(fdtd_engine* h);
int  fdtd_last_error(char* buf, size_t len);
/* Library version string and the device name (for reports). */
const char* fdtd_version(void);
int  fdtd_device_name(char* buf, size_t len);

/* ---- operator (set once, before the first fdtd_iterate) ---- */
/* Full coefficient arrays, each [3][Nx][Ny][Nz]. */
int  fdtd_set_coeffs(fdtd_engine* h, int dtype,
                     const void* vv, const void* vi, const void* ii, const void* iv);
/* Compressed coefficients: per-cell index [Nx][Ny][Nz] into a table of
 * n_entries rows of 12 values {vv0,vv1,vv2, vi0,vi1,vi2, ii0,ii1,ii2, iv0,iv1,iv2}. */
int  fdtd_set_coeffs_compressed(fdtd_engine* h, const uint32_t* index,
                                int dtype, const void* table, uint32_t n_entries);
/* Factored coefficients (no counterpart in the reference; for FP32, FP16
 * and BF16 handles, whose graded meshes defeat the plain compression): a
 * per-cell index [Nx][Ny][Nz] into at most 65536 rows of 12 floats
 * {vv0..2, m_E0..2, ii0..2, m_H0..2} and per-line mesh factors fv, fc
 * [3][Nx+Ny+Nz]; the engine uses vi = m_E * ((fv[n][x] * fv[n][Nx+y]) *
 * fv[n][Nx+Ny+z]) and iv likewise with fc, in float arithmetic in that order
 * (wop_factor builds all four).  The update kernels read 2 bytes of index per
 * cell instead of 24 bytes of coefficients. */
int  fdtd_set_coeffs_factored(fdtd_engine* h, const uint32_t* index, const float* table, uint32_t n_entries,
                              const float* fv, const float* fc);
/* Per-component storage scales for FP16/BF16 handles (volt, curr); stored
 * value = physical value / scale.  Ignored for FP32/FP64 handles. */
int  fdtd_set_storage_scale(fdtd_engine* h, const double volt_scale[3], const double curr_scale[3]);

int  fdtd_set_excitation(fdtd_engine* h, const fdtd_excitation* exc);

/* One UPML slab: lines start..start+num_lines-1 per axis; six coefficient
 * arrays, each [3][nx][ny][nz].  Slabs must not overlap. */
int  fdtd_add_upml(fdtd_engine* h, const uint32_t start[3], const uint32_t num_lines[3], int dtype,
                   const void* vv, const void* vvfo, const void* vvfn,
                   const void* ii, const void* iifo, const void* iifn);

/* One dispersive (Drude/Lorentz/Debye ADE) order: positions [3][count] and
 * coefficient lists [3][count].  Pass NULL for the lists of a disabled part;
 * flags bit0 = volt ADE on, bit1 = curr ADE on, bit2 = volt Lorentz on,
 * bit3 = curr Lorentz on (operator_ext_lorentzmaterial.h m_*_ADE_On). */
int  fdtd_add_ade(fdtd_engine* h, uint32_t count, const uint32_t* pos, int dtype,
                  const void* v_int, const void* v_ext, const void* v_lor,
                  const void* i_int, const void* i_ext, const void* i_lor, uint32_t flags);

/* Mur ABC face (Engine_Ext_Mur_ABC, engine_ext_mur_abc.cpp:83-168;
 * coefficients from Operator_Ext_Mur_ABC::BuildExtension,
 * operator_ext_mur_abc.cpp:117-180): face = 2*axis + top (x-lo 0 .. z-hi 5),
 * k_nyP/k_nyPP the first-order coefficients of the two tangential
 * components over the boundary plane, [N(axis+1)][N(axis+2)], start_ts the
 * timestep from which the face acts (wop_get_mur).  Hooks run in the
 * reference's extension order: pre before the ADE pre hook, post after the
 * TFSF injection, apply after the ADE apply and before the excitation. */
int  fdtd_add_mur(fdtd_engine* h, int face, int dtype, const void* k_nyP, const void* k_nyPP, uint32_t start_ts);
/* Lumped RLC cells (Engine_Ext_LumpedRLC, engine_ext_lumpedRLC.cpp:81-149;
 * lists of Operator_Ext_LumpedRLC, operator_ext_lumpedRLC.cpp:110-452):
 * pos [3][count], dir [count], coef [9][count] = ilv, i2v, vv2, vj1, vj2,
 * vvd, ib0, b1, b2 (wop_get_rlc).  Replaces any earlier list; count 0
 * clears it.  Each edge may appear once.  Pre hook after the Lorentz pre
 * hook, apply before the Lorentz apply. */
int  fdtd_set_rlc(fdtd_engine* h, uint32_t count, const uint32_t* pos, const uint32_t* dir, int dtype, const void* coef);
/* Steady-state detection for periodic excitations: Operator_Ext_SteadyState
 * (FDTD/extensions/operator_ext_steadystate.cpp:21-25,44-50,64-76, period in
 * seconds rounded to whole timesteps, one voltage probe per pos [3][count] /
 * dir [count] entry, at most 256) and Engine_Ext_SteadyState::Apply2Voltages
 * (engine_ext_steadystate.cpp:51-117).  Runs first among the voltage apply
 * hooks; every step records the probes, every period from the second on it
 * compares the last two periods and the domain energy (FP64).  period <= 0
 * removes it.  Set before the first timestep. */
int  fdtd_set_steady_state(fdtd_engine* h, double period, uint32_t count, const uint32_t* pos, const uint32_t* dir);
/* the period in timesteps, 0 when no detection is set */
int  fdtd_get_steady_state_period(fdtd_engine* h, uint32_t* period_ts);
/* Engine_Ext_SteadyState::GetLastDiff (engine_ext_steadystate.h:38): 1 until
 * the first evaluation; synchronises the engine's stream */
int  fdtd_steady_state_last_diff(fdtd_engine* h, double* diff);
/* Total-field/scattered-field injection list.  which = FDTD_VOLT (added after
 * the voltage update using the current signal) or FDTD_CURR (added after the
 * current update using the voltage signal).  Each entry adds
 *   (1-delta)*amp*s[L[delay]] + delta*amp*s[L[delay+1]]
 * to component comp at pos, L[n] being the delay lookup of engine_ext_tfsf.cpp:43-52. */
int  fdtd_set_tfsf(fdtd_engine* h, int which, uint32_t count, const uint32_t* pos,
                   const uint32_t* comp, const uint32_t* delay, const double* delta,
                   const double* amp, uint32_t max_delay);

/* ---- probes ---- */
/* Voltage: line from start to stop (one axis differs).  Current: rectangle
 * start..stop with normal norm_dir; inside_mask bits 0-2 = start inside per
 * axis, bits 3-5 = stop inside per axis (ProcessIntegral m_start_inside /
 * m_stop_inside), pass 0x3f for a probe inside the domain.
 * td_interval: time-domain sampling interval in timesteps (0 = none).
 * fd_interval: FD accumulation interval (0 = none) for n_freq frequencies. */
int  fdtd_add_probe(fdtd_engine* h, int kind, const uint32_t start[3], const uint32_t stop[3],
                    int norm_dir, uint32_t inside_mask, double weight,
                    uint32_t td_interval, uint32_t fd_interval,
                    const double* freqs, uint32_t n_freq, int* id);
int  fdtd_probe_num_samples(fdtd_engine* h, int id, uint32_t* n);
/* Copies min(n, max_n) samples (time in s, weighted value). */
int  fdtd_read_probe(fdtd_engine* h, int id, double* t, double* val, uint32_t max_n);
/* FD accumulators (weighted, FP64): re[n_freq], im[n_freq]. */
int  fdtd_read_probe_fd(fdtd_engine* h, int id, double* re, double* im);

/* ---- frequency-domain field boxes ---- */
/* Edge lengths in metres per axis, concatenated x, y, z ([Nx+Ny+Nz] each):
 * primary = Operator::GetEdgeLength(n, pos, false), dual = (..., true)
 * (FDTD/operator.cpp:151-207).  Needed before fdtd_add_dft_box. */
int  fdtd_set_edge_lengths(fdtd_engine* h, const double* primary, const double* dual);
/* Accumulate E (field = FDTD_VOLT) or H (FDTD_CURR) node-interpolated
 * fields of the mesh nodes start..stop (inclusive) every fd_interval steps
 * at n_freq (<= 64) frequencies: X += F(T) * exp(-2 pi i f T) * 2 * dT * fd_interval,
 * T = numTS*dT for E and (numTS+0.5)*dT for H.  Samples are taken at every
 * numTS divisible by fd_interval, numTS = 0 included. */
int  fdtd_add_dft_box(fdtd_engine* h, int field, const uint32_t start[3], const uint32_t stop[3],
                      const double* freqs, uint32_t n_freq, uint32_t fd_interval, int* id);
int  fdtd_dft_num_samples(fdtd_engine* h, int id, uint64_t* n);
/* Copies the accumulators, [n_freq][3][nx][ny][nz][2] doubles (re, im), z fastest. */
int  fdtd_read_dft(fdtd_engine* h, int id, double* out);
/* Time-domain field dump of a node box (ProcessFieldsTD::Process,
 * Common/processfields_td.cpp:54-91, over ProcessFields::CalcField,
 * Common/processfields.cpp:280-325, NODE_INTERPOLATE): at every timestep
 * divisible by `interval` (step 0 included) the interpolated E (field =
 * FDTD_VOLT) or H (FDTD_CURR) of nodes start..stop is rounded to float and
 * handed to `sink` as [3][nx][ny][nz] (z fastest) with the timestep and its
 * time (H at the dual time (ts + 1/2) dT).  The copy to the host overlaps the
 * following steps: a sample reaches the sink when its host slot is needed
 * again or before fdtd_iterate returns, always in timestep order, on the
 * thread that called fdtd_iterate; `field` is valid during the call only
 * and the sink must not call into the engine.  Needs fdtd_set_edge_lengths. */
typedef void (*fdtd_td_sink)(void* user, int id, uint32_t ts, double time, const float* field, uint64_t count);
int  fdtd_add_td_box(fdtd_engine* h, int field, const uint32_t start[3], const uint32_t stop[3], uint32_t interval,
                     fdtd_td_sink sink, void* user, int* id);
int  fdtd_td_num_samples(fdtd_engine* h, int id, uint64_t* n);

/* ---- dump types and interpolations of the field boxes ---- */
/* Dump types of ProcessFields (Common/processfields.h:43): E, H, current
 * density J = kappa E, rot H, flux densities D = epsilon E and B = mu H.  The
 * values are those of Engine_Interface_FDTD::GetRawField /
 * GetRawDualField (FDTD/engine_interface_fdtd.cpp:132-142,257-288). */
#define FDTD_DUMP_E    0
#define FDTD_DUMP_H    1
#define FDTD_DUMP_J    2
#define FDTD_DUMP_ROTH 3
#define FDTD_DUMP_D    4
#define FDTD_DUMP_B    5
/* Engine_Interface_Base::InterpolationType (Common/engine_interface_base.h:33):
 * NO_INTERPOLATION, NODE_INTERPOLATE, CELL_INTERPOLATE; the primary-grid
 * types (E, J, rot H, D) and the dual-grid types (H, B) each follow
 * GetRawInterpolatedField / GetRawInterpolatedDualField
 * (FDTD/engine_interface_fdtd.cpp:60-116,143-198). */
#define FDTD_INTERP_NONE 0
#define FDTD_INTERP_NODE 1
#define FDTD_INTERP_CELL 2
/* Averaged material of every edge as the operator stores it for dumps
 * (Operator::Calc_ECPos, FDTD/operator.cpp:1176-1188, m_epsR / m_kappa /
 * m_mueR): kind FDTD_MAT_EPS (absolute permittivity, F/m), FDTD_MAT_KAPPA
 * (S/m) or FDTD_MAT_MU (absolute permeability, H/m), float [3][Nx][Ny][Nz].
 * J dumps need kappa, D dumps epsilon and B dumps mu, set before the box is
 * added (the reference's SetMaterialStoreFlags, openems.cpp:617-624). */
#define FDTD_MAT_EPS   0
#define FDTD_MAT_KAPPA 1
#define FDTD_MAT_MU    2
int  fdtd_set_material(fdtd_engine* h, int kind, const float* values);
/* fdtd_add_dft_box / fdtd_add_td_box for any dump type and interpolation:
 * fdtd_add_dft_box(h, FDTD_VOLT, ...) is dump type E with node
 * interpolation and FDTD_CURR dump type H with node interpolation.  Only
 * H boxes use the dual time (ts + 1/2) dT (openems.cpp:595-600); J, rot H,
 * D and B are sampled at ts dT like E.  Values outside the ranges, or a
 * missing material, fail with FDTD_ERR_ARG / FDTD_ERR_STATE. */
int  fdtd_add_dft_dump_box(fdtd_engine* h, int dump_type, int interp, const uint32_t start[3], const uint32_t stop[3],
                           const double* freqs, uint32_t n_freq, uint32_t fd_interval, int* id);
int  fdtd_add_td_dump_box(fdtd_engine* h, int dump_type, int interp, const uint32_t start[3], const uint32_t stop[3],
                          uint32_t interval, fdtd_td_sink sink, void* user, int* id);

/* ---- time stepping ---- */
int  fdtd_iterate(fdtd_engine* h, uint32_t nsteps);
int  fdtd_synchronize(fdtd_engine* h);
int  fdtd_get_num_timesteps(fdtd_engine* h, uint32_t* numTS);
/* Tile order of the main update launches.  A tile is 256 vectors of
 * consecutive (y, z) cells of one plane.  band_tiles = 0: the launch sweeps
 * whole planes; n > 0: the tiles of a plane are cut into y bands of n tiles
 * and the launch sweeps band by band, each over all its planes, so the
 * neighbour plane a tile reads again is still in L2 where a whole plane
 * would not fit; -1 (default) = the engine's choice.  xcd = n > 0: the
 * eight XCDs take chunks of n consecutive tiles of every band-plane in turn
 * (XCD j chunks j, j + 8, ...; n >= band_tiles / 8 gives each XCD one
 * contiguous sub-band), so the neighbour rows and planes its blocks read
 * again were loaded by the same XCD; 0 off; -1 the engine's choice.
 * Results do not depend on the order, only speed does.  fdtd_tile_order
 * reports the order in use for the voltage update (band_tiles 0 = whole
 * planes, xcd = chunk width in use). */
int  fdtd_set_tile_order(fdtd_engine* h, int band_tiles, int xcd);
int  fdtd_tile_order(fdtd_engine* h, int* band_tiles, int* xcd);
/* UPML coefficients: the vector kernels can read per-cell uint16 row
 * indices into a table of distinct coefficient tuples (possible when each
 * side has at most 65536 distinct tuples) instead of the full slab arrays.
 * mode 1 = tables, 0 = full arrays, -1 = tables together with full main
 * coefficients or with the interior/shell split (default: a single launch
 * with compressed main coefficients loses more to the extra gathers than it
 * saves in bytes).  fdtd_pml_compressed reports whether the tables are in
 * use. */
int  fdtd_set_pml_compression(fdtd_engine* h, int mode);
int  fdtd_pml_compressed(fdtd_engine* h, int* active);
/* ADE coefficients (Lorentz/Drude orders): the ADE pre-update can read a
 * uint16 row index per entry into a table of the distinct coefficient
 * 9-tuples {int, ext, Lorentz} x 3 components of one field instead of the
 * per-entry lists (the entries of one material share a few tuples).  mode
 * 1 / -1 (default) = tables where a field has at most 65536 distinct
 * tuples, 0 = lists.  Applies to the orders added after the call.
 * fdtd_ade_compressed reports the form of order `order` per field.  Results
 * are identical. */
int  fdtd_set_ade_compression(fdtd_engine* h, int mode);
int  fdtd_ade_compressed(fdtd_engine* h, int order, int* volt, int* curr);
/* UPML shell launch (vector kernels): when the slabs form a shell around an
 * interior box (as the operator's UPML does), the interior can run in a
 * launch without UPML code and the shell in a launch of its own, instead of
 * one launch whose wavefronts branch into the UPML form.  mode 1 = split,
 * 0 = one launch, -1 (default) = the engine's choice: split when the
 * interior holds at least half of the cells, for FP16/BF16 and FP32 storage;
 * one launch for FP64.  Results are identical in every form.
 * fdtd_pml_split reports whether the split launches are in use. */
int  fdtd_set_pml_split(fdtd_engine* h, int mode);
int  fdtd_pml_split(fdtd_engine* h, int* active);
/* With the split: the interior launch's z range starts and ends at
 * multiples of `vectors` 16-byte vectors (rounded inwards; the shell launch
 * takes the cells in between), so the two launches share fewer partial
 * cache lines at the row ends.  1, 2, 4 or 8; -1 (default) = the engine's
 * choice.  Results are identical. */
int  fdtd_set_shell_align(fdtd_engine* h, int vectors);
/* Fused sparse hooks: the TFSF injection, the dispersive apply and the soft
 * excitation of a half step run as one launch that walks each touched field
 * location's hooks in the reference's order, and that launch also performs
 * the other field's ADE pre-update of the next half step (which no hook of
 * the current half step changes).  Results are identical to one launch per
 * hook.  mode 1 / -1 (default) = fused where no Mur ABC, lumped RLC or
 * steady-state hook sits between them, 0 = one launch per hook.
 * fdtd_hooks_fused reports whether the fused form is in use. */
int  fdtd_set_fuse_hooks(fdtd_engine* h, int mode);
int  fdtd_hooks_fused(fdtd_engine* h, int* active);
/* Use a captured HIP graph for the step loop (default on). */
int  fdtd_set_use_graph(fdtd_engine* h, int enable);
/* Kernel timing: while enabled, fdtd_iterate runs eagerly and records HIP
 * events around each main update kernel on the handle's stream.
 * fdtd_get_kernel_times returns the summed milliseconds and launch counts of
 * the voltage and current update kernels since the last reset. */
int  fdtd_set_kernel_timing(fdtd_engine* h, int enable);
int  fdtd_get_kernel_times(fdtd_engine* h, double* volt_ms, double* curr_ms,
                           uint32_t* volt_launches, uint32_t* curr_launches);

/* ---- field access ---- */
int  fdtd_calc_energy(fdtd_engine* h, double* energy);
/* fdtd_calc_energy without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc), which fdtd_calc_energy does on its first call per handle.  The
 * partial sums live in a library-owned static device array of 2 * 2048 + 2
 * doubles (32 KiB) per GPU, held until the process ends; the reduction uses
 * 2048 blocks for every domain size, so no domain exceeds it.  The result is
 * bitwise equal to fdtd_calc_energy's.  Every call on one GPU shares that
 * array: calls that overlap on one GPU (two handles on two threads or
 * streams) race and must be serialised by the caller. */
int  fdtd_calc_energy_noalloc(fdtd_engine* h, double* energy);
/* Per-component max |value| of the volt (FDTD_VOLT) or curr field in
 * physical units, reduced on the device.  No counterpart in the reference:
 * the storage-scale pilot of FP16/BF16 handles reads it instead of copying
 * whole fields to the host. */
int  fdtd_field_absmax(fdtd_engine* h, int which, double out[3]);
int  fdtd_get_field(fdtd_engine* h, int which, int dtype, void* out);
int  fdtd_set_field(fdtd_engine* h, int which, int dtype, const void* in);
int  fdtd_get_value(fdtd_engine* h, int which, uint32_t n, uint32_t x, uint32_t y, uint32_t z, double* v);
int  fdtd_set_value(fdtd_engine* h, int which, uint32_t n, uint32_t x, uint32_t y, uint32_t z, double v);

/* Streaming copy of `bytes` bytes, `reps` times, on a private stream:
 * achieved read+write bandwidth in GB/s.  Reference for the roofline and the
 * calibration kernel (k_copy) for HBM byte counters. */
int  fdtd_copy_bench(uint64_t bytes, int reps, double* gbps);

/* ---- batches of independent domains ----
 * A batch groups engine handles configured with the calls above (same
 * storage precision; operator, excitation, UPML and probes; all at the same
 * timestep) so that one launch per update phase advances every domain.
 * The handles stay owned by the caller and must outlive the batch; results
 * are read with the per-handle calls after fdtd_batch_synchronize.  While a
 * handle belongs to a batch it must not be stepped with fdtd_iterate.
 * No counterpart in the reference, which runs one simulation per process;
 * this is the batched-sweep pattern (parameter sweeps, per-port excitations). */
typedef struct fdtd_batch fdtd_batch;
int  fdtd_batch_create(fdtd_engine* const* engines, int count, fdtd_batch** out);
int  fdtd_batch_destroy(fdtd_batch* b);
/* threads per block of the batched update kernels: 64, 128 or 256 (tuning) */
int  fdtd_batch_set_block(fdtd_batch* b, int threads);
/* UPML launch form of the batch members: 1 / -1 (default) = every member
 * whose UPML slabs form a shell runs its interior in a launch without UPML
 * code and its shell in a shell launch shared by all such members; 0 = each
 * member's whole domain in the UPML form.  Results are identical. */
int  fdtd_batch_set_split(fdtd_batch* b, int mode);
int  fdtd_batch_set_use_graph(fdtd_batch* b, int enable);
int  fdtd_batch_iterate(fdtd_batch* b, uint32_t nsteps);
int  fdtd_batch_synchronize(fdtd_batch* b);
int  fdtd_batch_set_kernel_timing(fdtd_batch* b, int enable);
int  fdtd_batch_get_kernel_times(fdtd_batch* b, double* volt_ms, double* curr_ms, uint32_t* launches);

/* Device bytes held by the handle (fields, coefficients, extensions). */
int  fdtd_device_bytes(fdtd_engine* h, uint64_t* bytes);

#ifdef __cplusplus
}
#endif
#endif /* WIPEMS_FDTD_H */
