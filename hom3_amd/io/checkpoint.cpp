// HDF5 checkpoint / restart of solver states over the hom3g C ABI (see
// include/hom3g_io.h).  Host code; the device state is downloaded and
// uploaded through libhom3g.
#include <hdf5.h>

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
void h5(herr_t r, const char* what) {
  if (r < 0) throw Fail(HOM3G_ERR_STATE, std::string("HDF5: ") + what);
}
hid_t h5id(hid_t r, const char* what) {
  if (r < 0) throw Fail(HOM3G_ERR_STATE, std::string("HDF5: ") + what);
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
  int64_t ncells, first_ghost, nvars, nd, physics, step;
  double time, dt;
};

Info info_of(hom3g_ctx* ctx, int sid) {
  Info I{};
  int nd = 0, nv = 0, ph = 0;
  ck(hom3g_solver_info(ctx, sid, &nd, &nv, &ph));
  ck(hom3g_solver_counts(ctx, sid, &I.ncells, &I.first_ghost));
  ck(hom3g_solver_time(ctx, sid, &I.time, &I.dt, &I.step));
  I.nd = nd;
  I.nvars = nv;
  I.physics = ph;
  return I;
}

void attr_i64(hid_t g, const char* name, int64_t v) {
  const hid_t sp = h5id(H5Screate(H5S_SCALAR), "space");
  const hid_t a = h5id(H5Acreate2(g, name, H5T_NATIVE_INT64, sp, H5P_DEFAULT, H5P_DEFAULT), name);
  h5(H5Awrite(a, H5T_NATIVE_INT64, &v), name);
  H5Aclose(a);
  H5Sclose(sp);
}
void attr_f64(hid_t g, const char* name, double v) {
  const hid_t sp = h5id(H5Screate(H5S_SCALAR), "space");
  const hid_t a = h5id(H5Acreate2(g, name, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT), name);
  h5(H5Awrite(a, H5T_NATIVE_DOUBLE, &v), name);
  H5Aclose(a);
  H5Sclose(sp);
}
int64_t read_i64(hid_t g, const char* name) {
  int64_t v = 0;
  const hid_t a = h5id(H5Aopen(g, name, H5P_DEFAULT), name);
  h5(H5Aread(a, H5T_NATIVE_INT64, &v), name);
  H5Aclose(a);
  return v;
}
double read_f64(hid_t g, const char* name) {
  double v = 0;
  const hid_t a = h5id(H5Aopen(g, name, H5P_DEFAULT), name);
  h5(H5Aread(a, H5T_NATIVE_DOUBLE, &v), name);
  H5Aclose(a);
  return v;
}

}  // namespace

extern "C" {

const char* hom3g_io_last_error(void) { return g_err.c_str(); }

int hom3g_io_write_checkpoint(hom3g_ctx* ctx, const int* sids, int n, const char* path) {
  return guarded([&] {
    if (!ctx || !sids || n <= 0 || !path) throw Fail(HOM3G_ERR_ARG, "bad argument");
    const hid_t f = h5id(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), "create file");
    try {
      for (int k = 0; k < n; ++k) {
        const Info I = info_of(ctx, sids[k]);
        std::vector<double> q((size_t)I.nvars * I.ncells);
        ck(hom3g_solver_download_state(ctx, sids[k], q.data(), I.ncells));
        const std::string gname = "/solver" + std::to_string(k);
        const hid_t g = h5id(H5Gcreate2(f, gname.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "group");
        const hsize_t dims[2] = {(hsize_t)I.nvars, (hsize_t)I.ncells};
        const hid_t sp = h5id(H5Screate_simple(2, dims, nullptr), "space");
        const hid_t d = h5id(H5Dcreate2(g, "lhs", H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "lhs");
        h5(H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, q.data()), "write lhs");
        H5Dclose(d);
        H5Sclose(sp);
        attr_f64(g, "time", I.time);
        attr_f64(g, "dt", I.dt);
        attr_i64(g, "step", I.step);
        attr_i64(g, "ncells", I.ncells);
        attr_i64(g, "first_ghost", I.first_ghost);
        attr_i64(g, "nvars", I.nvars);
        attr_i64(g, "nd", I.nd);
        attr_i64(g, "physics", I.physics);
        H5Gclose(g);
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
      for (int k = 0; k < n; ++k) {
        const Info I = info_of(ctx, sids[k]);
        const std::string gname = "/solver" + std::to_string(k);
        const hid_t g = h5id(H5Gopen2(f, gname.c_str(), H5P_DEFAULT), "group");
        if (read_i64(g, "ncells") != I.ncells || read_i64(g, "first_ghost") != I.first_ghost ||
            read_i64(g, "nvars") != I.nvars || read_i64(g, "nd") != I.nd || read_i64(g, "physics") != I.physics) {
          H5Gclose(g);
          throw Fail(HOM3G_ERR_ARG, "checkpoint " + gname + " does not match the solver's topology");
        }
        std::vector<double> q((size_t)I.nvars * I.ncells);
        const hid_t d = h5id(H5Dopen2(g, "lhs", H5P_DEFAULT), "lhs");
        h5(H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, q.data()), "read lhs");
        H5Dclose(d);
        const double t = read_f64(g, "time"), dt = read_f64(g, "dt");
        const int64_t st = read_i64(g, "step");
        H5Gclose(g);
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
