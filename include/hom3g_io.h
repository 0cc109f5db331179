/* HDF5 checkpoint / restart of solver states (libhom3g_io.so, built on the
 * C ABI of libhom3g).
 *
 * The reference declares a "restart" property (solver.hpp:88-95, always
 * false in its tests) and an HDF5 file handler, io::file::HDF5File
 * (io/hdf5_file.hpp:40-385), whose conventions this file follows: arrays are
 * rank-1 datasets (H5LTmake_dataset) with attributes no_rows, no_cols (long
 * long) and order (int, 1 = column-major); single values are attributes
 * (long long, int or double).  Per solver k (in call order) the group
 * "/solver<k>" holds
 *   datasets  lhs       double    ncells x nvars (cells().lhs: internal
 *                                 cells, then ghosts; container.hpp:58-66)
 *             node_idx  long long ncells x 1 (invalid = INT64_MAX)
 *             bc_idx    int       ncells x 1
 *   attributes time, dt (double), step, first_ghost (long long), nd, nvars,
 *              physics (int);
 * the root carries no_solvers (long long).  Reading checks that the file's
 * cells, their order and boundary ids are the solver's, then restores lhs
 * and the clock.  The conventional extension is ".hdf5"
 * (hdf5::extension(), io/hdf5_file.hpp:31).  Status codes as in hom3g.h. */
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
