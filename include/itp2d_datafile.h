/*
 * itp2d_datafile.h -- C ABI of the HDF5 result file (the reference's
 * Datafile, datafile.hpp:40-123, datafile.cpp:32-467), in the reference's
 * layout so that its scripts and users' h5py/MATLAB readers work unchanged:
 *
 *   /states                            [slot][state] of {r, i} double arrays [sy][sx]
 *   /state_history                     [k] of {int step; int index}
 *   /final_energies                    [N] double
 *   /final_energy_standard_deviations  [N] double
 *   /energy_history                    [step][N] double
 *   /deviation_history                 [step][N] double
 *   /time_step_history                 [k] of {int step; double time_step} (offsets 0, 8)
 *   /potential_values                  scalar of a double array [sy][sx]
 *   /noise_data                        [k] double
 *   root attributes grid_sizex, grid_sizey, grid_delta and the run's
 *   parameters and results; a "description" string on the root and on
 *   every dataset.
 *
 * Host-only library (libitp2d_datafile.so, linked to the HDF5 C library);
 * independent of the GPU library. Every call returns ITP_OK (0) or a
 * negative status; itp_datafile_last_error gives the message of the last
 * failure on the calling thread.
 */
#ifndef ITP2D_DATAFILE_H
#define ITP2D_DATAFILE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct itp_datafile itp_datafile;

/* Datafile::Datafile (datafile.cpp:32-96): creates the file (an existing one
 * only when clobber is non-zero, else an error), the root description and
 * the grid attributes grid_sizex, grid_sizey (int) and grid_delta (double). */
int itp_datafile_create(const char* path, size_t sx, size_t sy, double dx, int clobber, itp_datafile** out);
/* Datafile::~Datafile: flushes and closes. */
void itp_datafile_close(itp_datafile* f);
/* Datafile::flush */
int itp_datafile_flush(itp_datafile* f);

/* Datafile::add_attribute overloads (datafile.cpp:416-467): a scalar root
 * attribute of type int, unsigned long, double or a fixed-length string.
 * An attribute that already exists is an error, as in the reference. */
int itp_datafile_attr_int(itp_datafile* f, const char* name, int value);
int itp_datafile_attr_ulong(itp_datafile* f, const char* name, unsigned long value);
int itp_datafile_attr_double(itp_datafile* f, const char* name, double value);
int itp_datafile_attr_string(itp_datafile* f, const char* name, const char* value);

/* Datafile::write_stateset (datafile.cpp:246-275): appends one slot to
 * /states holding N states (interleaved complex doubles [N][sy][sx][2]), in
 * the order order[0..N) when order is non-null (the sorted save of
 * ITPSystem::save_states, itpsystem.cpp:374-388), and the pair (step, slot)
 * to /state_history. */
int itp_datafile_write_states(itp_datafile* f, int step, const double* states, size_t N, const size_t* order);
/* Datafile::write_time_step_history (datafile.cpp:277-297): appends (step, eps). */
int itp_datafile_write_time_step(itp_datafile* f, int step, double eps);
/* Datafile::write_energy_history / write_deviation_history (datafile.cpp:299-316,
 * 338-355): row `step` of /energy_history or /deviation_history (which = 0 or
 * 1), N values; the dataset grows to step + 1 rows. */
int itp_datafile_write_history_row(itp_datafile* f, int which, size_t step, const double* values, size_t N);
/* Datafile::write_energies / write_energy_standard_deviations (datafile.cpp:
 * 324-336, 363-375): /final_energies (which = 0) or
 * /final_energy_standard_deviations (which = 1), N values. */
int itp_datafile_write_final(itp_datafile* f, int which, const double* values, size_t N);
/* Datafile::write_potential (datafile.cpp:377-388): the sy x sx values. */
int itp_datafile_write_potential(itp_datafile* f, const double* V);
/* Datafile::write_noise_realization (datafile.cpp:390-405): n values. */
int itp_datafile_write_noise(itp_datafile* f, const double* data, size_t n);

/* Reads a double dataset (any rank, or a scalar double array) of a file
 * into out (capacity cap values); *count receives its number of values
 * (out = NULL: only the count). For checking a written file; no
 * counterpart in the reference. */
int itp_datafile_read_doubles(const char* path, const char* dataset, double* out, size_t cap, size_t* count);

const char* itp_datafile_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
