// Checkpoint files: savefile / Restart! (src/integrator.jl:208-288) and readMHDFlows
// (src/utils/IC.jl:245-257) with the reference's dataset names, in a small self-describing
// container (libhdf5 is not available to this library).
//
// Layout (little-endian): "MHDFLOWS", uint32 version = 1, uint32 count, then count entries of
// {char name[48]; uint32 dtype; uint32 ndim; uint64 dims[4]; uint64 offset; uint64 nbytes},
// then the data blocks at 64-byte aligned offsets.  dims are in Julia (column-major) order,
// first dimension fastest, so a real field is (nx, ny, nz) and sol is (nkr, ny, nz, nf).
// dtype: 1 Float32, 2 Float64, 3 Complex{Float32}, 4 Complex{Float64}, 5 Int64.
//
// Datasets written by every save (the reference's savefile): i/j/k_velocity, i/j/k_mag_field
// (MHD) = prob.vars, and time = clock.t.  A bit-exact save adds sol (the spectral state),
// clock_step, clock_dt and cfl_maxima (the maxima getCFL! uses next), so a restart from it
// continues exactly as the uninterrupted run.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mhd_engine.hpp"

namespace mhd {

namespace {

constexpr char kMagic[8] = {'M', 'H', 'D', 'F', 'L', 'O', 'W', 'S'};
enum DType : uint32_t { DT_F32 = 1, DT_F64 = 2, DT_C32 = 3, DT_C64 = 4, DT_I64 = 5 };

struct Entry {
  char name[48];
  uint32_t dtype;
  uint32_t ndim;
  uint64_t dims[4];
  uint64_t offset;
  uint64_t nbytes;
};
static_assert(sizeof(Entry) == 104, "checkpoint entry layout");

size_t dtype_bytes(uint32_t d) {
  switch (d) {
    case DT_F32: return 4;
    case DT_F64: return 8;
    case DT_C32: return 8;
    case DT_C64: return 16;
    case DT_I64: return 8;
    default: return 0;
  }
}

struct Dataset {
  std::string name;
  uint32_t dtype;
  std::vector<uint64_t> dims;
  std::vector<char> data;
};

const char* kFieldNames[6] = {"i_velocity", "j_velocity", "k_velocity", "i_mag_field", "j_mag_field", "k_mag_field"};

int write_file(const char* path, const std::vector<Dataset>& ds) {
  std::vector<Entry> ents(ds.size());
  uint64_t off = 16 + sizeof(Entry) * ds.size();
  off = (off + 63) / 64 * 64;
  for (size_t i = 0; i < ds.size(); ++i) {
    std::memset(&ents[i], 0, sizeof(Entry));
    std::strncpy(ents[i].name, ds[i].name.c_str(), sizeof(ents[i].name) - 1);
    ents[i].dtype = ds[i].dtype;
    ents[i].ndim = (uint32_t)ds[i].dims.size();
    for (size_t d = 0; d < ds[i].dims.size(); ++d) ents[i].dims[d] = ds[i].dims[d];
    ents[i].offset = off;
    ents[i].nbytes = ds[i].data.size();
    off = (off + ents[i].nbytes + 63) / 64 * 64;
  }
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(E_INVALID, std::string("cannot open ") + path + " for writing");
  bool ok = std::fwrite(kMagic, 1, 8, f) == 8;
  const uint32_t ver = 1, cnt = (uint32_t)ds.size();
  ok = ok && std::fwrite(&ver, 4, 1, f) == 1 && std::fwrite(&cnt, 4, 1, f) == 1;
  ok = ok && std::fwrite(ents.data(), sizeof(Entry), ents.size(), f) == ents.size();
  for (size_t i = 0; ok && i < ds.size(); ++i) {
    ok = std::fseek(f, (long)ents[i].offset, SEEK_SET) == 0;
    ok = ok && (ds[i].data.empty() || std::fwrite(ds[i].data.data(), 1, ds[i].data.size(), f) == ds[i].data.size());
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(E_INVALID, std::string("writing ") + path + " failed");
  return 0;
}

int read_file(const char* path, std::vector<Entry>* ents, FILE** fp) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(E_INVALID, std::string("cannot open ") + path);
  char magic[8];
  uint32_t ver = 0, cnt = 0;
  bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMagic, 8) == 0;
  ok = ok && std::fread(&ver, 4, 1, f) == 1 && std::fread(&cnt, 4, 1, f) == 1 && ver == 1 && cnt < 4096;
  if (ok) {
    ents->resize(cnt);
    ok = std::fread(ents->data(), sizeof(Entry), cnt, f) == cnt;
  }
  if (!ok) {
    std::fclose(f);
    return fail(E_INVALID, std::string(path) + " is not an MHDFlows checkpoint file");
  }
  *fp = f;
  return 0;
}

const Entry* find(const std::vector<Entry>& ents, const char* name) {
  for (const Entry& e : ents)
    if (std::strncmp(e.name, name, sizeof(e.name)) == 0) return &e;
  return nullptr;
}

int read_block(FILE* f, const Entry& e, void* dst) {
  if (std::fseek(f, (long)e.offset, SEEK_SET) != 0 || std::fread(dst, 1, e.nbytes, f) != e.nbytes)
    return fail(E_INVALID, std::string("reading dataset ") + e.name + " failed");
  return 0;
}

template <class T>
void put_scalar(std::vector<Dataset>* ds, const char* name, uint32_t dt, T v) {
  Dataset d{name, dt, {}, std::vector<char>(sizeof(T))};
  std::memcpy(d.data.data(), &v, sizeof(T));
  ds->push_back(std::move(d));
}

}  // namespace

int save_checkpoint(EngineBase* e, int run, const char* path, int exact) {
  if (!path || !*path) return fail(E_INVALID, "checkpoint path is empty");
  const Geom& g = e->geom();
  const int nf = e->nfields();
  const bool dbl = e->precision_bytes() == 8;
  const size_t rb = dbl ? 8 : 4;
  std::vector<Dataset> ds;
  for (int f = 0; f < nf; ++f) {  // savefile writes prob.vars (integrator.jl:263-277)
    Dataset d{kFieldNames[f], dbl ? DT_F64 : DT_F32, {(uint64_t)g.nx, (uint64_t)g.ny, (uint64_t)g.nz},
              std::vector<char>(rb * (size_t)g.nr)};
    MHD_TRY(e->get_vars(run, f, d.data.data()));
    ds.push_back(std::move(d));
  }
  RunClock c;
  MHD_TRY(e->clock_io(run, &c, 0));
  if (dbl) put_scalar<double>(&ds, "time", DT_F64, c.t);  // write(fw, "time", prob.clock.t) (:286), in T
  else put_scalar<float>(&ds, "time", DT_F32, (float)c.t);
  if (exact) {
    Dataset d{"sol", dbl ? DT_C64 : DT_C32, {(uint64_t)g.nkr, (uint64_t)g.ny, (uint64_t)g.nz, (uint64_t)nf},
              std::vector<char>(2 * rb * (size_t)g.nc * nf)};
    MHD_TRY(e->get_sol(run, d.data.data()));
    ds.push_back(std::move(d));
    put_scalar<int64_t>(&ds, "clock_step", DT_I64, c.step);
    if (dbl) put_scalar<double>(&ds, "clock_dt", DT_F64, c.dt);
    else put_scalar<float>(&ds, "clock_dt", DT_F32, (float)c.dt);
    Dataset m{"cfl_maxima", DT_F64, {2}, std::vector<char>(16)};
    const double mx[2] = {c.maxu2, c.maxb2};
    std::memcpy(m.data.data(), mx, 16);
    ds.push_back(std::move(m));
  }
  return write_file(path, ds);
}

int restart_checkpoint(EngineBase* e, int run, const char* path, int exact) {
  if (!path || !*path) return fail(E_INVALID, "checkpoint path is empty");
  std::vector<Entry> ents;
  FILE* f = nullptr;
  MHD_TRY(read_file(path, &ents, &f));
  struct Closer {
    FILE* f;
    ~Closer() { std::fclose(f); }
  } closer{f};
  const Geom& g = e->geom();
  const int nf = e->nfields();
  const bool dbl = e->precision_bytes() == 8;
  const size_t rb = dbl ? 8 : 4;
  RunClock c;
  MHD_TRY(e->clock_io(run, &c, 0));
  const Entry* te = find(ents, "time");
  if (!te || te->ndim != 0 || (te->dtype != DT_F32 && te->dtype != DT_F64))
    return fail(E_INVALID, std::string(path) + ": no scalar dataset 'time'");
  double t = 0;
  if (te->dtype == DT_F64) {
    MHD_TRY(read_block(f, *te, &t));
  } else {
    float tf = 0;
    MHD_TRY(read_block(f, *te, &tf));
    t = tf;
  }
  if (exact) {
    const Entry* se = find(ents, "sol");
    const Entry* ke = find(ents, "clock_step");
    const Entry* de = find(ents, "clock_dt");
    const Entry* me = find(ents, "cfl_maxima");
    if (!se || !ke || !de || !me) return fail(E_INVALID, std::string(path) + " holds no bit-exact state (save it with exact = 1)");
    const uint64_t want[4] = {(uint64_t)g.nkr, (uint64_t)g.ny, (uint64_t)g.nz, (uint64_t)nf};
    if (se->ndim != 4 || se->dtype != (dbl ? DT_C64 : DT_C32) || std::memcmp(se->dims, want, sizeof(want)) != 0 ||
        se->nbytes != 2 * rb * (uint64_t)g.nc * nf)
      return fail(E_INVALID, std::string(path) + ": sol does not match this problem's grid, fields or precision");
    std::vector<char> sol(se->nbytes);
    MHD_TRY(read_block(f, *se, sol.data()));
    int64_t step = 0;
    double mx[2] = {0, 0}, dt = 0;
    MHD_TRY(read_block(f, *ke, &step));
    if (de->dtype == DT_F64) {
      MHD_TRY(read_block(f, *de, &dt));
    } else {
      float d32 = 0;
      MHD_TRY(read_block(f, *de, &d32));
      dt = d32;
    }
    MHD_TRY(read_block(f, *me, mx));
    MHD_TRY(e->set_sol(run, sol.data()));
    c.t = t;
    c.dt = dt;
    c.step = step;
    c.maxu2 = mx[0];
    c.maxb2 = mx[1];
    c.nan_flag = 0;
    return e->clock_io(run, &c, 1);
  }
  // Restart! (integrator.jl:208-257): real fields -> vars, R2C into sol; clock.t = time
  std::vector<std::vector<char>> fields(nf);
  std::vector<const void*> ptrs(nf, nullptr);
  for (int k = 0; k < nf; ++k) {
    const Entry* fe = find(ents, kFieldNames[k]);
    if (!fe) return fail(E_INVALID, std::string(path) + ": no dataset " + kFieldNames[k]);
    const uint64_t want[3] = {(uint64_t)g.nx, (uint64_t)g.ny, (uint64_t)g.nz};
    if (fe->ndim != 3 || std::memcmp(fe->dims, want, sizeof(want)) != 0 ||
        (fe->dtype != DT_F32 && fe->dtype != DT_F64) || fe->nbytes != dtype_bytes(fe->dtype) * (uint64_t)g.nr)
      return fail(E_INVALID, std::string(path) + ": " + kFieldNames[k] + " does not match this problem's grid");
    std::vector<char> raw(fe->nbytes);
    MHD_TRY(read_block(f, *fe, raw.data()));
    fields[k].resize(rb * (size_t)g.nr);
    // copyto! converts to the problem's type (IC.jl / utils.jl:172-176)
    if ((fe->dtype == DT_F64) == dbl) {
      fields[k].swap(raw);
    } else if (dbl) {
      const float* s = (const float*)raw.data();
      double* d = (double*)fields[k].data();
      for (int64_t i = 0; i < g.nr; ++i) d[i] = s[i];
    } else {
      const double* s = (const double*)raw.data();
      float* d = (float*)fields[k].data();
      for (int64_t i = 0; i < g.nr; ++i) d[i] = (float)s[i];
    }
    ptrs[k] = fields[k].data();
  }
  MHD_TRY(e->set_ic_real(run, ptrs.data()));
  MHD_TRY(e->clock_io(run, &c, 0));
  c.t = dbl ? t : (double)(float)t;
  return e->clock_io(run, &c, 1);
}

}  // namespace mhd
