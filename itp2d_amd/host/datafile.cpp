// HDF5 result file in the reference's layout (datafile.cpp:32-467), written
// with the HDF5 C API. Datasets are created on first use, extendible along
// their first dimension (unlimited maximum extents), chunked and deflated.
// Chunks: one state per chunk for /states (each element is a whole state
// array, as in the reference); one history row per chunk; 1024 values for
// the one-dimensional datasets.
#include <hdf5.h>

#include <cstddef>
#include <cstring>
#include <string>
#include <vector>

#include "itp2d_datafile.h"

namespace {

constexpr int kOK = 0, kInvalid = -1, kIO = -10;

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// Closes an HDF5 identifier at scope exit.
struct Id {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  Id() = default;
  Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  Id(const Id&) = delete;
  Id& operator=(const Id&) = delete;
  ~Id() {
    if (id >= 0 && close) close(id);
  }
  bool ok() const { return id >= 0; }
  operator hid_t() const { return id; }
};

struct StateHistoryPair {
  int step;
  int index;
};
struct TimeStepPair {
  int step;
  double time_step;
};

}  // namespace

struct itp_datafile {
  hid_t file = -1;
  hsize_t sx = 0, sy = 0;
  hid_t complex_t = -1, state_t = -1, potential_t = -1, sh_t = -1, ts_t = -1;
};

namespace {

int set_description(hid_t obj, const char* text) {
  Id st(H5Tcopy(H5T_C_S1), H5Tclose);
  const size_t len = std::strlen(text);
  if (!st.ok() || (len && H5Tset_size(st, len) < 0)) return fail(kIO, "string type");
  Id sp(H5Screate(H5S_SCALAR), H5Sclose);
  Id a(H5Acreate2(obj, "description", st, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (!a.ok() || H5Awrite(a, st, text) < 0) return fail(kIO, "description attribute");
  return kOK;
}

// Opens a dataset, creating it (rank-dimensional, extent 0 along the
// extendible dimensions, chunk dims `chunk`) if it does not exist.
hid_t open_or_create(itp_datafile* f, const char* name, hid_t type, int rank, const hsize_t* chunk,
                     const char* description, int deflate = 6) {
  if (H5Lexists(f->file, name, H5P_DEFAULT) > 0) return H5Dopen2(f->file, name, H5P_DEFAULT);
  hsize_t zero[2] = {0, 0}, unl[2] = {H5S_UNLIMITED, H5S_UNLIMITED};
  Id sp(H5Screate_simple(rank, zero, unl), H5Sclose);
  Id pl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
  if (!sp.ok() || !pl.ok() || H5Pset_chunk(pl, rank, chunk) < 0) return -1;
  if (deflate > 0 && H5Pset_deflate(pl, deflate) < 0) return -1;
  hid_t d = H5Dcreate2(f->file, name, type, sp, H5P_DEFAULT, pl, H5P_DEFAULT);
  if (d < 0) return -1;
  if (set_description(d, description) != kOK) {
    H5Dclose(d);
    return -1;
  }
  return d;
}

int dims_of(hid_t d, int rank, hsize_t* dims) {
  Id sp(H5Dget_space(d), H5Sclose);
  if (!sp.ok() || H5Sget_simple_extent_ndims(sp) != rank) return kIO;
  H5Sget_simple_extent_dims(sp, dims, nullptr);
  return kOK;
}

// Writes `count` elements of `mem_t` from buf into d at the hyperslab
// (start, count) of its (already extended) file space.
int write_slab(hid_t d, hid_t mem_t, int rank, const hsize_t* start, const hsize_t* count, const void* buf) {
  Id fs(H5Dget_space(d), H5Sclose);
  if (!fs.ok() || H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, nullptr, count, nullptr) < 0) return kIO;
  Id ms(H5Screate_simple(rank, count, nullptr), H5Sclose);
  if (!ms.ok() || H5Dwrite(d, mem_t, ms, fs, H5P_DEFAULT, buf) < 0) return kIO;
  return kOK;
}

template <class T>
int add_attr(itp_datafile* f, const char* name, hid_t type, const T* value) {
  if (!f || !name) return fail(kInvalid, "invalid argument");
  if (H5Aexists(f->file, name) > 0) return fail(kInvalid, std::string("attribute exists: ") + name);
  Id sp(H5Screate(H5S_SCALAR), H5Sclose);
  Id a(H5Acreate2(f->file, name, type, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (!a.ok() || H5Awrite(a, type, value) < 0) return fail(kIO, std::string("attribute ") + name);
  return kOK;
}

const char* kRootText =
    "Results of an itp2d imaginary-time propagation run. Energies, lengths and the magnetic field are in "
    "Hartree atomic units (field strength in the SI-based atomic unit); timings are in seconds. Each "
    "dataset carries its own description.";
const char* kStatesText =
    "Saved states, indexed [slot][state]. A slot is one save of the whole set (e.g. one per iteration, "
    "or the final set); within a slot the states may be in order of increasing energy. Each state is a "
    "[sy][sx] array of complex wave-function values on the grid.";
const char* kStateHistText =
    "One (step, index) pair per saved slot of /states: the iteration step at which the set was saved "
    "and the slot index it went to.";
const char* kEnergiesText = "The last computed energy of each state.";
const char* kDeviationsText = "The last computed energy standard deviation of each state.";
const char* kEnergyHistText =
    "Energy of each state at each iteration step, indexed [step][state] (states in order of energy).";
const char* kDeviationHistText =
    "Energy standard deviation of each state at each iteration step, indexed [step][state] (states in "
    "order of energy).";
const char* kTimeStepText =
    "One (step, time_step) pair per change of the imaginary time step: the iteration step from which "
    "the new value applies, and the value.";
const char* kPotentialText = "Values of the external potential on the grid, [sy][sx].";
const char* kNoiseText = "Raw description of the random potential noise realization, if one was used.";

}  // namespace

extern "C" {

const char* itp_datafile_last_error(void) { return g_err.c_str(); }

int itp_datafile_create(const char* path, size_t sx, size_t sy, double dx, int clobber, itp_datafile** out) {
  if (!path || !out || sx == 0 || sy == 0) return fail(kInvalid, "invalid argument");
  *out = nullptr;
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);  // errors come back as status codes
  hid_t file = H5Fcreate(path, clobber ? H5F_ACC_TRUNC : H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT);
  if (file < 0)
    return fail(kIO, std::string("cannot create datafile '") + path +
                         "' (an existing file needs clobber; the directory must exist)");
  itp_datafile* f = new itp_datafile;
  f->file = file;
  f->sx = sx;
  f->sy = sy;
  f->complex_t = H5Tcreate(H5T_COMPOUND, 2 * sizeof(double));
  H5Tinsert(f->complex_t, "r", 0, H5T_NATIVE_DOUBLE);
  H5Tinsert(f->complex_t, "i", sizeof(double), H5T_NATIVE_DOUBLE);
  const hsize_t sdims[2] = {sy, sx};
  f->state_t = H5Tarray_create2(f->complex_t, 2, sdims);
  f->potential_t = H5Tarray_create2(H5T_NATIVE_DOUBLE, 2, sdims);
  f->sh_t = H5Tcreate(H5T_COMPOUND, sizeof(StateHistoryPair));
  H5Tinsert(f->sh_t, "step", offsetof(StateHistoryPair, step), H5T_NATIVE_INT);
  H5Tinsert(f->sh_t, "index", offsetof(StateHistoryPair, index), H5T_NATIVE_INT);
  f->ts_t = H5Tcreate(H5T_COMPOUND, sizeof(TimeStepPair));
  H5Tinsert(f->ts_t, "step", offsetof(TimeStepPair, step), H5T_NATIVE_INT);
  H5Tinsert(f->ts_t, "time_step", offsetof(TimeStepPair, time_step), H5T_NATIVE_DOUBLE);
  int rc = kOK;
  {
    Id root(H5Gopen2(file, "/", H5P_DEFAULT), H5Gclose);
    if (!root.ok() || f->state_t < 0 || f->potential_t < 0) rc = fail(kIO, "datafile types");
    if (!rc) rc = set_description(root, kRootText);
  }
  const int isx = (int)sx, isy = (int)sy;
  if (!rc) rc = itp_datafile_attr_int(f, "grid_sizex", isx);
  if (!rc) rc = itp_datafile_attr_int(f, "grid_sizey", isy);
  if (!rc) rc = itp_datafile_attr_double(f, "grid_delta", dx);
  if (rc) {
    itp_datafile_close(f);
    return rc;
  }
  *out = f;
  return kOK;
}

void itp_datafile_close(itp_datafile* f) {
  if (!f) return;
  for (hid_t t : {f->complex_t, f->state_t, f->potential_t, f->sh_t, f->ts_t})
    if (t >= 0) H5Tclose(t);
  if (f->file >= 0) H5Fclose(f->file);
  delete f;
}

int itp_datafile_flush(itp_datafile* f) {
  if (!f) return fail(kInvalid, "invalid argument");
  return H5Fflush(f->file, H5F_SCOPE_GLOBAL) < 0 ? fail(kIO, "flush") : kOK;
}

int itp_datafile_attr_int(itp_datafile* f, const char* name, int value) {
  return add_attr(f, name, H5T_NATIVE_INT, &value);
}
int itp_datafile_attr_ulong(itp_datafile* f, const char* name, unsigned long value) {
  return add_attr(f, name, H5T_NATIVE_ULONG, &value);
}
int itp_datafile_attr_double(itp_datafile* f, const char* name, double value) {
  return add_attr(f, name, H5T_NATIVE_DOUBLE, &value);
}
int itp_datafile_attr_string(itp_datafile* f, const char* name, const char* value) {
  if (!value) return fail(kInvalid, "invalid argument");
  Id st(H5Tcopy(H5T_C_S1), H5Tclose);
  const size_t len = std::strlen(value);
  if (!st.ok() || (len && H5Tset_size(st, len) < 0)) return fail(kIO, "string type");
  return add_attr(f, name, (hid_t)st, value);
}

int itp_datafile_write_states(itp_datafile* f, int step, const double* states, size_t N, const size_t* order) {
  if (!f || !states || N == 0) return fail(kInvalid, "invalid argument");
  if (order)
    for (size_t m = 0; m < N; m++)
      if (order[m] >= N) return fail(kInvalid, "sort order index out of range");
  const hsize_t chunk2[2] = {1, 1}, chunk1[1] = {64};
  Id d(open_or_create(f, "/states", f->state_t, 2, chunk2, kStatesText), H5Dclose);
  Id h(open_or_create(f, "/state_history", f->sh_t, 1, chunk1, kStateHistText), H5Dclose);
  if (!d.ok() || !h.ok()) return fail(kIO, "states datasets");
  hsize_t cur[2], hcur[1];
  if (dims_of(d, 2, cur) || dims_of(h, 1, hcur)) return fail(kIO, "states extent");
  const hsize_t slot = cur[0];
  const hsize_t ext[2] = {slot + 1, cur[1] > N ? cur[1] : N};
  const hsize_t hext[1] = {hcur[0] + 1};
  if (H5Dset_extent(d, ext) < 0 || H5Dset_extent(h, hext) < 0) return fail(kIO, "extend states");
  const StateHistoryPair pair = {step, (int)slot};
  const hsize_t hstart[1] = {hcur[0]}, one[1] = {1};
  if (write_slab(h, f->sh_t, 1, hstart, one, &pair)) return fail(kIO, "write state_history");
  const size_t per = (size_t)f->sx * f->sy * 2;
  if (!order) {
    const hsize_t start[2] = {slot, 0}, count[2] = {1, N};
    if (write_slab(d, f->state_t, 2, start, count, states)) return fail(kIO, "write states");
  } else {
    for (size_t m = 0; m < N; m++) {
      const hsize_t start[2] = {slot, m}, count[2] = {1, 1};
      if (write_slab(d, f->state_t, 2, start, count, states + order[m] * per)) return fail(kIO, "write states");
    }
  }
  return kOK;
}

int itp_datafile_write_time_step(itp_datafile* f, int step, double eps) {
  if (!f) return fail(kInvalid, "invalid argument");
  const hsize_t chunk[1] = {64};
  Id d(open_or_create(f, "/time_step_history", f->ts_t, 1, chunk, kTimeStepText), H5Dclose);
  hsize_t cur[1];
  if (!d.ok() || dims_of(d, 1, cur)) return fail(kIO, "time_step_history dataset");
  const hsize_t ext[1] = {cur[0] + 1}, one[1] = {1};
  TimeStepPair p;
  std::memset(&p, 0, sizeof(p));
  p.step = step;
  p.time_step = eps;
  if (H5Dset_extent(d, ext) < 0 || write_slab(d, f->ts_t, 1, cur, one, &p)) return fail(kIO, "write time step");
  return kOK;
}

int itp_datafile_write_history_row(itp_datafile* f, int which, size_t step, const double* values, size_t N) {
  if (!f || !values || N == 0 || (which != 0 && which != 1)) return fail(kInvalid, "invalid argument");
  const hsize_t chunk[2] = {1, N};
  Id d(open_or_create(f, which ? "/deviation_history" : "/energy_history", H5T_NATIVE_DOUBLE, 2, chunk,
                      which ? kDeviationHistText : kEnergyHistText),
       H5Dclose);
  hsize_t cur[2];
  if (!d.ok() || dims_of(d, 2, cur)) return fail(kIO, "history dataset");
  const hsize_t ext[2] = {cur[0] > step + 1 ? cur[0] : step + 1, cur[1] > N ? cur[1] : N};
  const hsize_t start[2] = {step, 0}, count[2] = {1, N};
  if (H5Dset_extent(d, ext) < 0 || write_slab(d, H5T_NATIVE_DOUBLE, 2, start, count, values))
    return fail(kIO, "write history row");
  return kOK;
}

int itp_datafile_write_final(itp_datafile* f, int which, const double* values, size_t N) {
  if (!f || !values || N == 0 || (which != 0 && which != 1)) return fail(kInvalid, "invalid argument");
  const hsize_t chunk[1] = {N < 1024 ? N : 1024};
  Id d(open_or_create(f, which ? "/final_energy_standard_deviations" : "/final_energies", H5T_NATIVE_DOUBLE, 1,
                      chunk, which ? kDeviationsText : kEnergiesText),
       H5Dclose);
  const hsize_t ext[1] = {N}, start[1] = {0};
  if (!d.ok() || H5Dset_extent(d, ext) < 0 || write_slab(d, H5T_NATIVE_DOUBLE, 1, start, ext, values))
    return fail(kIO, "write final values");
  return kOK;
}

int itp_datafile_write_potential(itp_datafile* f, const double* V) {
  if (!f || !V) return fail(kInvalid, "invalid argument");
  if (H5Lexists(f->file, "/potential_values", H5P_DEFAULT) > 0) return fail(kInvalid, "potential already written");
  Id sp(H5Screate(H5S_SCALAR), H5Sclose);
  Id d(H5Dcreate2(f->file, "/potential_values", f->potential_t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
       H5Dclose);
  if (!d.ok() || set_description(d, kPotentialText) ||
      H5Dwrite(d, f->potential_t, H5S_ALL, H5S_ALL, H5P_DEFAULT, V) < 0)
    return fail(kIO, "write potential");
  return kOK;
}

int itp_datafile_write_noise(itp_datafile* f, const double* data, size_t n) {
  if (!f || (!data && n)) return fail(kInvalid, "invalid argument");
  const hsize_t chunk[1] = {64};
  Id d(open_or_create(f, "/noise_data", H5T_NATIVE_DOUBLE, 1, chunk, kNoiseText), H5Dclose);
  if (!d.ok()) return fail(kIO, "noise dataset");
  const hsize_t ext[1] = {n}, start[1] = {0};
  if (H5Dset_extent(d, ext) < 0 || (n && write_slab(d, H5T_NATIVE_DOUBLE, 1, start, ext, data)))
    return fail(kIO, "write noise");
  return kOK;
}

int itp_datafile_read_doubles(const char* path, const char* dataset, double* out, size_t cap, size_t* count) {
  if (!path || !dataset || !count) return fail(kInvalid, "invalid argument");
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  Id file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!file.ok()) return fail(kIO, std::string("cannot open ") + path);
  Id d(H5Dopen2(file, dataset, H5P_DEFAULT), H5Dclose);
  if (!d.ok()) return fail(kIO, std::string("no dataset ") + dataset);
  Id sp(H5Dget_space(d), H5Sclose);
  Id t(H5Dget_type(d), H5Tclose);
  // doubles per element: 1 for double data, the array's size for array types
  size_t per = 1;
  if (H5Tget_class(t) == H5T_ARRAY) {
    Id base(H5Tget_super(t), H5Tclose);
    if (H5Tget_class(base) == H5T_COMPOUND) return fail(kInvalid, "complex data: not a double dataset");
    per = H5Tget_size(t) / sizeof(double);
  } else if (H5Tget_class(t) != H5T_FLOAT) {
    return fail(kInvalid, "not a double dataset");
  }
  const hssize_t n = H5Sget_simple_extent_npoints(sp);
  if (n < 0) return fail(kIO, "extent");
  *count = (size_t)n * per;
  if (!out || *count == 0) return kOK;  // size query
  if (cap < *count) return fail(kInvalid, "output too small");
  if (H5Dread(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0) return fail(kIO, "read");
  return kOK;
}

}  // extern "C"
