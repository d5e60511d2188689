/* wipems_dump.h -- HDF5 field-dump files in the reference's layout.
 *
 * One file per dump box, as ProcessFields writes them
 * (Common/processfields.cpp:126-158, tools/hdf5_file_writer.cpp):
 *
 *   /                      attribute openEMS_HDF5_version = 0.2 (double)
 *   /Mesh/x, /Mesh/y, /Mesh/z
 *                          float mesh lines of the box, line * scaling
 *                          evaluated in float (HDF5_File_Writer::WriteRectMesh,
 *                          tools/hdf5_file_writer.cpp:107-197)
 *   /FieldData/TD/<ts>     one dataset per time sample, named by the timestep
 *                          zero-padded to 8 digits, float [3][nz][ny][nx],
 *                          attribute time (float) (ProcessFieldsTD::Process,
 *                          Common/processfields_td.cpp:54-91; WriteVectorField,
 *                          tools/hdf5_file_writer.cpp:286-304)
 *   /FieldData/FD          attribute frequency (double [n_freq])
 *                          (ProcessFieldsFD::InitProcess,
 *                          Common/processfields_fd.cpp:55-62)
 *   /FieldData/FD/f<k>_real, f<k>_imag
 *                          float [3][nz][ny][nx] per frequency, attribute
 *                          frequency (float) (ProcessFieldsFD::DumpFDData,
 *                          Common/processfields_fd.cpp:202-220)
 *
 * Field arguments use the engine's layout [3][nx][ny][nz] (z fastest); the
 * writer transposes to the file's [3][nz][ny][nx] (x fastest), as the
 * reference's WriteVectorField does.
 *
 * Every function returns 0 on success and -1 on failure;
 * wdump_last_error() returns the message of the last failure on the calling
 * thread.  Host code only (HDF5 C library); no GPU calls.
 */
#ifndef WIPEMS_DUMP_H
#define WIPEMS_DUMP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wdump_file wdump_file;

int wdump_last_error(char* buf, size_t len);

/* HDF5_File_Writer(filename) + WriteRectMesh + the version attribute: creates
 * (truncates) the file.  lines_*: mesh lines of the box in drawing units;
 * scaling: metres per drawing unit (Operator::GetGridDelta). */
int wdump_create(const char* path, const double* lines_x, uint32_t nx, const double* lines_y, uint32_t ny,
                 const double* lines_z, uint32_t nz, double scaling, wdump_file** out);

/* One time sample: /FieldData/TD/<ts> with attribute time = (float)time.
 * field: [3][nx][ny][nz] float. */
int wdump_write_td(wdump_file* w, uint32_t ts, double time, const float* field);

/* Frequency-domain result: the /FieldData/FD frequency attribute and, per
 * frequency k, f<k>_real / f<k>_imag with attribute frequency.
 * acc: [n_freq][3][nx][ny][nz] complex values as (re, im) double pairs,
 * written as float like the reference's complex<float> accumulators. */
int wdump_write_fd(wdump_file* w, uint32_t n_freq, const double* freqs, const double* acc);

/* SAR dumps of an E-field box with cell interpolation (ProcessFieldsSAR,
 * Common/processfields_sar.cpp:179-335): kappa, density and volume are the
 * cells' centre conductivity, mass density and volume [nx][ny][nz]
 * (wop_get_cell_materials).  raw = 0: local SAR (dump type 20), per
 * frequency /FieldData/FD/f<k> float [nz][ny][nx] with P = 0.5 * sum_n kappa
 * |E_n|^2, SAR = P / density (0 where the density is 0), attributes
 * frequency and power = sum P * volume (SAR_Calculation::CalcLocalSAR /
 * CalcSARPower, tools/sar_calculation.cpp:162-230).  raw = 1 (type 29): the
 * E accumulators as f<k>_real / f<k>_imag and /CellData/Conductivity,
 * /Density, /Volume.  The /FieldData/FD frequency attribute either way.
 * Mass-averaged SAR (types 21, 22) is not built. */
int wdump_write_sar(wdump_file* w, uint32_t n_freq, const double* freqs, const double* acc, const float* kappa,
                    const float* density, const float* volume, int raw);

int wdump_close(wdump_file* w);

#ifdef __cplusplus
}
#endif

#endif
