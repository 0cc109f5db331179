/* HDF5 checkpoint / restart of solver states (libhom3g_io.so, built on the
 * C ABI of libhom3g).
 *
 * The reference declares a "restart" property (solver.hpp:88-95, always
 * false in its tests) and an HDF5 file handler (io/hdf5_file.hpp:40-385:
 * datasets written from / read into memory, row- or column-major).  This
 * library writes, per solver, the group "/solver<k>" with
 *   dataset "lhs"      float64 [nvars][ncells]  (cells().lhs, internal cells
 *                                                then ghosts, column-major
 *                                                in the reference's sense)
 *   attributes time, dt (float64), step, ncells, first_ghost, nvars, nd,
 *              physics (int64)
 * and restores them with the same checks (a file written for another
 * topology is rejected).  Status codes as in hom3g.h. */
#ifndef HOM3G_IO_H_
#define HOM3G_IO_H_
#include <stdint.h>
#include "hom3g.h"
#ifdef __cplusplus
extern "C" {
#endif

int hom3g_io_write_checkpoint(hom3g_ctx* ctx, const int* solver_ids, int n, const char* path);
int hom3g_io_read_checkpoint(hom3g_ctx* ctx, const int* solver_ids, int n, const char* path);
const char* hom3g_io_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* HOM3G_IO_H_ */
