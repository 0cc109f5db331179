// HDF5 checkpoint / restart of solver states over the hom3g C ABI (see
// include/hom3g_io.h).  Host code; the device state is downloaded and
// uploaded through libhom3g.  The file follows the conventions of the
// reference's io::file::HDF5File (io/hdf5_file.hpp:40-385): every array is a
// rank-1 dataset made with H5LTmake_dataset carrying the attributes
// "no_rows", "no_cols" (Ind) and "order" (SInd, 0 row-major, 1 col-major,
// :144-158, 270-289), and attributes are single values written with
// H5LTset_attribute_long_long / _int / _double for Ind / SInd / Num
// (:195-250).
#include <hdf5.h>
#include <hdf5_hl.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "hom3g_io.h"

namespace {

thread_local std::string g_err;

struct Fail : std::runtime_error {
  int code;
  Fail(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void ck(int rc) {
  if (rc != HOM3G_OK) throw Fail(rc, hom3g_last_error());
}
void h5(herr_t r, const std::string& what) {
  if (r < 0) throw Fail(HOM3G_ERR_STATE, "HDF5: " + what);
}
hid_t h5id(hid_t r, const std::string& what) {
  if (r < 0) throw Fail(HOM3G_ERR_STATE, "HDF5: " + what);
  return r;
}

template <class F>
int guarded(F&& f) {
  try {
    f();
    return HOM3G_OK;
  } catch (const Fail& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return HOM3G_ERR_STATE;
  }
}

struct Info {
  int64_t ncells, first_ghost, step;
  int nvars, nd, physics;
  double time, dt;
};

Info info_of(hom3g_ctx* ctx, int sid) {
  Info I{};
  ck(hom3g_solver_info(ctx, sid, &I.nd, &I.nvars, &I.physics));
  ck(hom3g_solver_counts(ctx, sid, &I.ncells, &I.first_ghost));
  ck(hom3g_solver_time(ctx, sid, &I.time, &I.dt, &I.step));
  return I;
}

// attribute maps of HDF5File (io/hdf5_file.hpp:195-250)
void set_attr(hid_t f, const std::string& obj, const char* name, long long v) {
  h5(H5LTset_attribute_long_long(f, obj.c_str(), name, &v, 1), obj + " " + name);
}
void set_attr(hid_t f, const std::string& obj, const char* name, int v) {
  h5(H5LTset_attribute_int(f, obj.c_str(), name, &v, 1), obj + " " + name);
}
void set_attr(hid_t f, const std::string& obj, const char* name, double v) {
  h5(H5LTset_attribute_double(f, obj.c_str(), name, &v, 1), obj + " " + name);
}
long long get_ll(hid_t f, const std::string& obj, const char* name) {
  long long v = 0;
  h5(H5LTget_attribute_long_long(f, obj.c_str(), name, &v), obj + " " + name);
  return v;
}
int get_int(hid_t f, const std::string& obj, const char* name) {
  int v = 0;
  h5(H5LTget_attribute_int(f, obj.c_str(), name, &v), obj + " " + name);
  return v;
}
double get_dbl(hid_t f, const std::string& obj, const char* name) {
  double v = 0;
  h5(H5LTget_attribute_double(f, obj.c_str(), name, &v), obj + " " + name);
  return v;
}

// write_array_from_memory_ (io/hdf5_file.hpp:144-158): rank 1, no_rows x
// no_cols values, column-major (order 1)
template <class T>
void write_array(hid_t f, const std::string& name, const T* mem, long long no_rows, long long no_cols, hid_t type) {
  const hsize_t dim[1] = {(hsize_t)(no_rows * no_cols)};
  h5(H5LTmake_dataset(f, name.c_str(), 1, dim, type, mem), name);
  set_attr(f, name, "no_rows", no_rows);
  set_attr(f, name, "no_cols", no_cols);
  set_attr(f, name, "order", 1);
}

void check_dims(hid_t f, const std::string& name, long long rows, long long cols) {
  if (get_ll(f, name, "no_rows") != rows || get_ll(f, name, "no_cols") != cols || get_int(f, name, "order") != 1)
    throw Fail(HOM3G_ERR_ARG, name + ": dimensions or order do not match the solver");
}

}  // namespace

extern "C" {

const char* hom3g_io_last_error(void) { return g_err.c_str(); }

int hom3g_io_write_checkpoint(hom3g_ctx* ctx, const int* sids, int n, const char* path) {
  return guarded([&] {
    if (!ctx || !sids || n <= 0 || !path) throw Fail(HOM3G_ERR_ARG, "bad argument");
    const hid_t f = h5id(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), "create file");
    try {
      set_attr(f, "/", "no_solvers", (long long)n);
      for (int k = 0; k < n; ++k) {
        const Info I = info_of(ctx, sids[k]);
        std::vector<double> q((size_t)I.nvars * I.ncells);
        std::vector<int64_t> node(I.ncells);
        std::vector<int32_t> bc(I.ncells);
        ck(hom3g_solver_download_state(ctx, sids[k], q.data(), I.ncells));
        ck(hom3g_solver_topology(ctx, sids[k], nullptr, node.data(), bc.data(), nullptr, nullptr, nullptr));
        const std::string g = "/solver" + std::to_string(k);
        h5(H5Gclose(h5id(H5Gcreate2(f, g.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), g)), g);
        // cells().lhs (container.hpp:58-66): ncells x nvars, column-major
        write_array(f, g + "/lhs", q.data(), I.ncells, I.nvars, H5T_NATIVE_DOUBLE);
        write_array(f, g + "/node_idx", reinterpret_cast<const long long*>(node.data()), I.ncells, 1,
                    H5T_NATIVE_LLONG);
        write_array(f, g + "/bc_idx", bc.data(), I.ncells, 1, H5T_NATIVE_INT);
        set_attr(f, g, "time", I.time);
        set_attr(f, g, "dt", I.dt);
        set_attr(f, g, "step", (long long)I.step);
        set_attr(f, g, "first_ghost", (long long)I.first_ghost);
        set_attr(f, g, "nd", I.nd);
        set_attr(f, g, "nvars", I.nvars);
        set_attr(f, g, "physics", I.physics);
      }
    } catch (...) {
      H5Fclose(f);
      throw;
    }
    h5(H5Fclose(f), "close file");
  });
}

int hom3g_io_read_checkpoint(hom3g_ctx* ctx, const int* sids, int n, const char* path) {
  return guarded([&] {
    if (!ctx || !sids || n <= 0 || !path) throw Fail(HOM3G_ERR_ARG, "bad argument");
    const hid_t f = h5id(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), "open file");
    try {
      if (get_ll(f, "/", "no_solvers") != n) throw Fail(HOM3G_ERR_ARG, "checkpoint holds another number of solvers");
      for (int k = 0; k < n; ++k) {
        const Info I = info_of(ctx, sids[k]);
        const std::string g = "/solver" + std::to_string(k);
        if (get_int(f, g, "nd") != I.nd || get_int(f, g, "nvars") != I.nvars || get_int(f, g, "physics") != I.physics ||
            get_ll(f, g, "first_ghost") != I.first_ghost)
          throw Fail(HOM3G_ERR_ARG, "checkpoint " + g + " does not match the solver");
        check_dims(f, g + "/lhs", I.ncells, I.nvars);
        check_dims(f, g + "/node_idx", I.ncells, 1);
        check_dims(f, g + "/bc_idx", I.ncells, 1);
        // the same cells in the same order (topology of the solver)
        std::vector<int64_t> node(I.ncells), node_f(I.ncells);
        std::vector<int32_t> bc(I.ncells), bc_f(I.ncells);
        ck(hom3g_solver_topology(ctx, sids[k], nullptr, node.data(), bc.data(), nullptr, nullptr, nullptr));
        h5(H5LTread_dataset(f, (g + "/node_idx").c_str(), H5T_NATIVE_LLONG, node_f.data()), "node_idx");
        h5(H5LTread_dataset(f, (g + "/bc_idx").c_str(), H5T_NATIVE_INT, bc_f.data()), "bc_idx");
        if (node != node_f || bc != bc_f) throw Fail(HOM3G_ERR_ARG, "checkpoint " + g + " holds another topology");
        std::vector<double> q((size_t)I.nvars * I.ncells);
        h5(H5LTread_dataset(f, (g + "/lhs").c_str(), H5T_NATIVE_DOUBLE, q.data()), "lhs");
        const double t = get_dbl(f, g, "time"), dt = get_dbl(f, g, "dt");
        const long long st = get_ll(f, g, "step");
        ck(hom3g_solver_upload_state(ctx, sids[k], q.data(), I.ncells));
        ck(hom3g_solver_reset_time(ctx, sids[k], t, dt, st));
      }
    } catch (...) {
      H5Fclose(f);
      throw;
    }
    h5(H5Fclose(f), "close file");
  });
}

}  // extern "C"
