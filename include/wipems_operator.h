/*
 * wipems_operator.h -- C ABI of the host-side operator builder.
 *
 * This library plays the part of openEMS's Operator for the scenarios this
 * project runs (it replaces CSXCAD geometry with boxes and spheres).  It
 * builds the equivalent-circuit coefficients and the extension tables that
 * the engine ABI (wipems_fdtd.h) consumes:
 *
 *   wop_build         Operator::CalcECOperator (FDTD/operator.cpp:973-1084):
 *                     Calc_EC / AverageMatQuarterCell (:1406-1518) / Calc_ECPos (:1176-1224),
 *                     CalcTimestep_Var3 (:1897-1971), Calc_ECOperatorPos (:943-971),
 *                     ApplyElectricBC (:1086-1125), CalcPEC (:1973-2020),
 *                     Calc_LumpedElements (:1535-1712), ApplyMagneticBC (:1127-1174),
 *                     then the extensions' BuildExtension in creation order:
 *                     Operator_Ext_Excitation (operator_ext_excitation.cpp:101-285),
 *                     Operator_Ext_UPML (operator_ext_upml.cpp:77-466),
 *                     Operator_Ext_LorentzMaterial (operator_ext_lorentzmaterial.cpp:117-438),
 *                     and the excitation signal (FDTD/excitation.cpp:96-177).
 *   wop_compress      Operator_SSE_Compressed::CompressOperator
 *                     (FDTD/operator_sse_compressed.cpp:120-176), per cell instead
 *                     of per 4-lane vector.
 *
 * The build runs either with FDTD_FLOAT = float semantics (every EC value and
 * coefficient rounded to float where the reference stores float: prec 0) or
 * fully in double (prec 1, the FP64 oracle operator).
 *
 * Status: 0 = OK, negative = error; wop_last_error() gives the message.
 */
#ifndef WIPEMS_OPERATOR_H
#define WIPEMS_OPERATOR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wop_operator wop_operator;

/* boundary condition codes (openems.cpp BoundaryCond): 0 PEC, 1 PMC, 2 MUR, 3 PML, -1 none */
#define WOP_BC_PEC  0
#define WOP_BC_PMC  1
#define WOP_BC_MUR  2
#define WOP_BC_PML  3

/* excitation signal types (FDTD/excitation.h) */
#define WOP_EXC_GAUSS   0
#define WOP_EXC_SINE    1
#define WOP_EXC_DIRAC   2
#define WOP_EXC_STEP    3

int  wop_create(wop_operator** out);
int  wop_destroy(wop_operator* op);
int  wop_last_error(char* buf, size_t len);

/* mesh lines in drawing units; unit = metres per drawing unit (gridDelta) */
int  wop_set_mesh(wop_operator* op, int axis, const double* lines, uint32_t n);
int  wop_set_grid_delta(wop_operator* op, double unit);
int  wop_set_boundary(wop_operator* op, const int bc[6], const uint32_t pml_size[6]);
int  wop_set_timestep_factor(wop_operator* op, double factor);
/* Mur ABC (boundary code WOP_BC_MUR) phase velocity of one face
 * (face = 2*axis + top), 0 = the material's own (openems.cpp:749-754,
 * Operator_Ext_Mur_ABC::SetPhaseVelocity) */
int  wop_set_mur_phase_velocity(wop_operator* op, int face, double v_phase);
/* optional forced timestep (openems.cpp TimeStep): 0 = calculate */
int  wop_set_timestep(wop_operator* op, double dT);
int  wop_set_threads(wop_operator* op, int nthreads);
/* 1: the streamed build prints a progress line to stderr every 64 x planes
 * (openEMS' verbose level, openems.h:120 SetVerboseLevel) */
int  wop_set_verbose(wop_operator* op, int level);
/* 1 (default): the builds copy a cell's EC, Var3 timestep estimate and
 * coefficients from the cell below it in z where they are provably the same
 * (no primitive face near either cell, the mesh around them shifted bit for
 * bit); 0: every cell is evaluated.  The result is identical either way.  No
 * counterpart in the reference, whose Calc_EC_Range evaluates every cell
 * (FDTD/operator.cpp:1784-1809). */
int  wop_set_row_reuse(wop_operator* op, int on);

/* properties (CSXCAD stand-ins); each returns its id */
int  wop_add_material(wop_operator* op, double eps_r, double kappa, double mue_r, double sigma, int* id);
int  wop_add_metal(wop_operator* op, int* id);
/* Drude/Lorentz material of order 1: plasma frequency (Hz), relaxation time (s),
 * Lorentz pole frequency (Hz, 0 = Drude) for the electric response. */
int  wop_add_lorentz_material(wop_operator* op, double eps_r, double kappa, double mue_r, double sigma,
                              double eps_plasma_freq, double eps_relax_time, double eps_lor_pole_freq, int* id);
/* lumped parallel RC (R or C <= 0 means not given), direction ny, metal caps */
int  wop_add_lumped_element(wop_operator* op, int ny, double R, double C, int caps, int* id);
/* excitation property: type 0 soft E, 1 hard E, 2 soft H, 3 hard H; value vector; delay (s) */
/* lumped RLC element (CSPropLumpedElement with an inductance or of series
 * type; built by Operator_Ext_LumpedRLC, operator_ext_lumpedRLC.cpp:110-452):
 * type 0 parallel, 1 series; R, L, C < 0 = unset */
int  wop_add_lumped_rlc(wop_operator* op, int ny, int type, double R, double L, double C, int caps, int* id);
int  wop_add_excitation(wop_operator* op, int type, const double value[3], double delay, int* id);
int  wop_add_box(wop_operator* op, int prop, const double start[3], const double stop[3], int priority);
/* plane-wave (TFSF) excitation on the faces of a box (CSPropExcitation type 10
 * with one box primitive): propagation direction, E vector, and the frequency
 * for the numerical phase velocity (<= 0: c0).  One per operator. */
int  wop_set_plane_wave(wop_operator* op, const double start[3], const double stop[3], const double prop_dir[3],
                        const double e_vec[3], double freq);
int  wop_add_sphere(wop_operator* op, int prop, const double center[3], double radius, int priority);

/* excitation time signal */
int  wop_set_excite_signal(wop_operator* op, int type, double f0, double fc);

/* build: prec 0 = float storage semantics, 1 = double; nr_ts = max timesteps
 * (caps the signal length as Excitation::CalcGaussianPulsExcitation does) */
int  wop_build(wop_operator* op, int prec, uint32_t nr_ts);

/* results */
int  wop_get_num_lines(wop_operator* op, uint32_t n[3]);
/* Operator::GetEdgeLength(n, pos, dual) in metres for every line of each axis,
 * concatenated x, y, z: out[Nx+Ny+Nz] (FDTD/operator.cpp:151-207). */
int  wop_get_edge_lengths(wop_operator* op, int dual, double* out);
int  wop_get_timestep(wop_operator* op, double* dT);
/* Nyquist number of the excitation (Excitation::GetNyquistNum) and the
 * step of maximum excitation amplitude (GetMaxExcitationTimestep) */
int  wop_get_excite_info(wop_operator* op, uint32_t* length, uint32_t* nyquist, uint32_t* max_step);
int  wop_get_signal(wop_operator* op, double* volt, double* curr);
/* Excitation::GetSignalPeriod (FDTD/excitation.h:74; 1/f0 for a sine, else 0)
 * and the timestep count int(period/dT) at which the excitation and TFSF
 * hooks wrap the signal (engine_ext_excitation.cpp:43-45) */
int  wop_get_signal_period(wop_operator* op, double* period, uint32_t* period_ts);
/* coefficients as dense [3][Nx][Ny][Nz] arrays in dtype (0 f32, 1 f64) */
int  wop_get_coeffs(wop_operator* op, int dtype, void* vv, void* vi, void* ii, void* iv);
/* EC values after Calc_EC (before the operator build), [3][N] each, double */
int  wop_get_ec(wop_operator* op, double* C, double* G, double* L, double* R);
/* Material storage for field dumps (Operator_Base::SetMaterialStoreFlags,
 * Common/operator_base.cpp:83, with Operator::Calc_ECPos storing the
 * averaged EffMat[type] of every edge, FDTD/operator.cpp:1176-1188): type 0
 * absolute permittivity, 1 conductivity kappa, 2 absolute permeability.
 * Set before the build (full or streamed); wop_get_material then copies the
 * values, rounded to float as the reference stores them, [3][Nx][Ny][Nz]. */
int  wop_set_material_store_flags(wop_operator* op, int type, int val);
/* Mass density of a material property (kg/m^3; CSPropMaterial's Density),
 * for SAR. */
int  wop_set_density(wop_operator* op, int prop, double density);
/* Per mesh node of the box start..stop (inclusive), [nx][ny][nz] float: the
 * conductivity and density of the material at the cell's centre (the cell
 * from that node to the next lines; no material: 0) and the cell volume in
 * m^3, as ProcessFieldsSAR::DumpFDData gathers them
 * (Common/processfields_sar.cpp:200-230).  Any output may be null. */
int  wop_get_cell_materials(wop_operator* op, const uint32_t start[3], const uint32_t stop[3], float* kappa, float* density,
                            float* volume);
int  wop_get_material(wop_operator* op, int type, float* out);
/* compression of the coefficients: returns the number of unique 12-tuples and
 * (if index/table are non-null) the per-cell index and the table [n][12] */
int  wop_compress(wop_operator* op, uint32_t* n_entries, uint32_t* index, double* table, uint32_t max_entries);
/* Factored form of the coefficients for float engines with 2-byte field
 * storage (no counterpart in the reference, whose compression,
 * FDTD/operator_sse_compressed.cpp:120-176, dedups {vv,vi,ii,iv} as they
 * are): vi = m_E * ((fv[n][x] * fv[n][Nx+y]) * fv[n][Nx+Ny+z]) and
 * iv = m_H * ((fc[n][x] * fc[n][Nx+y]) * fc[n][Nx+Ny+z]) in float arithmetic,
 * fv/fc [3][Nx+Ny+Nz] the mesh factors of the EC formulas and m rounded to
 * `bits` (8..23) significant bits; index[Nx*Ny*Nz] (caller-owned) addresses
 * table[max_entries][12] floats {vv0..2, m_E0..2, ii0..2, m_H0..2}.
 * Returns -2 when more than max_entries tuples exist; max_rel_err receives
 * the largest |rebuilt - built| / |built| over nonzero vi/iv. */
int  wop_factor(wop_operator* op, int bits, uint32_t max_entries, uint32_t* index, float* table, uint32_t* n_entries,
                float* fv, float* fc, double* max_rel_err);
/* Float build for large domains that never holds full per-cell arrays
 * (Operator::CalcECOperator + Operator_SSE_Compressed::CompressOperator,
 * FDTD/operator.cpp:973-1084, FDTD/operator_sse_compressed.cpp:120-176):
 * x planes are built one at a time and deduplicated into index[Nx*Ny*Nz]
 * (caller-owned) as they are finished.  Same index and table as wop_build
 * followed by wop_compress.  No TFSF plane wave or Lorentz materials.
 * wop_get_streamed_table copies the table [n_entries][12]. */
int  wop_build_streamed(wop_operator* op, uint32_t nr_ts, uint32_t* index, uint32_t max_entries, uint32_t* n_entries);
int  wop_get_streamed_table(wop_operator* op, double* table);

int  wop_get_excitation_counts(wop_operator* op, uint32_t* volt_count, uint32_t* curr_count);
int  wop_get_excitation(wop_operator* op, uint32_t* volt_pos, uint32_t* volt_dir, double* volt_amp, uint32_t* volt_delay,
                        uint32_t* curr_pos, uint32_t* curr_dir, double* curr_amp, uint32_t* curr_delay);

/* TFSF injection lists (Operator_Ext_TFSF::BuildExtension, operator_ext_tfsf.cpp:82-410),
 * flattened in the order Engine_Ext_TFSF applies them: which = 0 voltage list
 * (after the voltage update), 1 current list.  pos is [3][count]. */
int  wop_get_tfsf_info(wop_operator* op, uint32_t* volt_count, uint32_t* curr_count, uint32_t* max_delay);
int  wop_get_tfsf(wop_operator* op, int which, uint32_t* pos, uint32_t* comp, uint32_t* delay, double* delta, double* amp);

/* Mur ABC faces after a build (Operator_Ext_Mur_ABC, operator_ext_mur_abc.cpp:117-180):
 * per face the coefficients of its two tangential components over the
 * boundary plane, [N(axis+1)][N(axis+2)], and the timestep from which the
 * boundary acts (Engine_Ext_Mur_ABC's delay for excitations on the plane,
 * engine_ext_mur_abc.cpp:42-58), for fdtd_add_mur */
int  wop_get_num_mur(wop_operator* op, uint32_t* n);
int  wop_get_mur(wop_operator* op, uint32_t k, int* face, int dtype, void* k_nyP, void* k_nyPP, uint32_t* start_ts);
/* lumped RLC cells after a build: pos [3][count], dir [count], coef [9][count]
 * = ilv, i2v, vv2, vj1, vj2, vvd, ib0, b1, b2 (Operator_Ext_LumpedRLC's lists),
 * for fdtd_set_rlc */
int  wop_get_rlc_count(wop_operator* op, uint32_t* count);
int  wop_get_rlc(wop_operator* op, uint32_t* pos, uint32_t* dir, int dtype, void* coef);
int  wop_get_num_upml(wop_operator* op, uint32_t* n);
int  wop_get_upml_range(wop_operator* op, uint32_t k, uint32_t start[3], uint32_t num_lines[3]);
int  wop_get_upml(wop_operator* op, uint32_t k, int dtype, void* vv, void* vvfo, void* vvfn, void* ii, void* iifo, void* iifn);

int  wop_get_num_ade(wop_operator* op, uint32_t* orders);
int  wop_get_ade_info(wop_operator* op, uint32_t order, uint32_t* count, uint32_t* flags);
int  wop_get_ade(wop_operator* op, uint32_t order, uint32_t* pos, int dtype, void* v_int, void* v_ext, void* v_lor,
                 void* i_int, void* i_ext, void* i_lor);

/* snapping helpers (Operator::SnapToMeshLine, FDTD/operator.cpp:244-273) */
int  wop_snap_to_mesh_line(wop_operator* op, int axis, double coord, int dual, uint32_t* idx, int* inside);

#ifdef __cplusplus
}
#endif
#endif /* WIPEMS_OPERATOR_H */
