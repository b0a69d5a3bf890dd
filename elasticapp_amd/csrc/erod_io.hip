// Checkpoints of a device batch: everything a batch holds, written to one
// binary file and read back into a new batch.
//
// The reference archives a rod block as its variable arrays plus its ghost
// node buffer (elastica/Systems/Block/Serialize.hpp:37-62,
// elastica/Systems/CosseratRods/Serialize.hpp:41-72).  Here the file also
// carries what lives beside the arrays on the device path (rod clocks, hook
// descriptors, per-rod dt and damping, active lengths, step budgets, the
// environment), so a run resumed from it continues bit for bit.
//
// Layout (little-endian, no padding):
//   "ERODCKP1"  u32 version (3)  u32 damped  u64 n_rods  u64 element_slots
//   u64 n_elems[n_rods]  u64 n_active[n_rods]  u32 steps_left[n_rods]  u8 budgets
//   f64 rod_time[n_rods]
//   per rod: f64 dt, gamma_f, gamma_t  u32 flags  f64 clamp 0 x[3], q[9]  f64 force[3]
//            i32 ramp  f64 ramp_start, ramp_interval
//            u32 clamp_idx[2]  f64 clamp 1 x[3], q[9]
//   (version 1 and 2 files, without the last line, load with clamp 0 at index
//   0 and no clamp 1)
//   i32 gravity  f64 g[3]  i32 contact  f64 k, c, mu, ct, z0, offset  f64 n[3]
//   (version 1 files, without the plane normal n, load with n = (0, 0, 1))
//   u32 n_fields; per field: u32 id  u32 ncomp  u64 count  f64 data[ncomp * count]
//   "ERODEND1"
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "erod_common.hpp"

using namespace erod;

namespace {
  constexpr char kMagic[9] = "ERODCKP1";
  constexpr char kEnd[9] = "ERODEND1";
  constexpr uint32_t kVersion = 3;

  struct File {
    std::FILE* f = nullptr;
    ~File() {
      if (f) std::fclose(f);
    }
  };

  template <typename T>
  bool put(std::FILE* f, const T& v) { return std::fwrite(&v, sizeof(T), 1, f) == 1; }
  template <typename T>
  bool put_n(std::FILE* f, const T* v, size_t n) { return n == 0 || std::fwrite(v, sizeof(T), n, f) == n; }
  template <typename T>
  bool get(std::FILE* f, T& v) { return std::fread(&v, sizeof(T), 1, f) == 1; }
  template <typename T>
  bool get_n(std::FILE* f, T* v, size_t n) { return n == 0 || std::fread(v, sizeof(T), n, f) == n; }

  int io_error(const std::string& what) {
    set_error(what);
    return EROD_ERR_INVALID;
  }
}  // namespace

extern "C" {

int erod_save(erod_batch* b, const char* path) {
  if (!b || !path) return io_error("erod_save: bad argument");
  std::vector<double> t(b->n_rods);
  if (int rc = erod_get_rod_time(b, t.data())) return rc;
  File F;
  F.f = std::fopen(path, "wb");
  if (!F.f) return io_error(std::string("erod_save: cannot open ") + path);
  std::FILE* f = F.f;
  bool ok = put_n(f, kMagic, 8) && put(f, kVersion) && put(f, uint32_t(b->damped)) && put(f, b->n_rods) && put(f, b->E);
  ok = ok && put_n(f, b->n_elems.data(), b->n_rods) && put_n(f, b->n_active.data(), b->n_rods) &&
       put_n(f, b->steps_left.data(), b->n_rods) && put(f, uint8_t(b->budgets)) && put_n(f, t.data(), b->n_rods);
  for (uint64_t r = 0; ok && r < b->n_rods; ++r) {
    const RodInfo& ri = b->rods[r];
    ok = put(f, ri.dt) && put(f, ri.gamma_f) && put(f, ri.gamma_t) && put(f, ri.flags) && put_n(f, ri.clamp[0].x, 3) &&
         put_n(f, ri.clamp[0].q, 9) && put_n(f, ri.force, 3) && put(f, int32_t(ri.ramp)) && put(f, ri.ramp_start) &&
         put(f, ri.ramp_interval) && put_n(f, ri.clamp_idx, 2) && put_n(f, ri.clamp[1].x, 3) && put_n(f, ri.clamp[1].q, 9);
  }
  const Env& e = b->env;
  ok = ok && put(f, int32_t(e.gravity)) && put_n(f, e.g, 3) && put(f, int32_t(e.contact)) && put(f, e.k) && put(f, e.c) &&
       put(f, e.mu) && put(f, e.ct) && put(f, e.z0) && put(f, e.offset) && put_n(f, e.n, 3) &&
       put(f, uint32_t(EROD_N_FIELDS));
  std::vector<double> buf;
  for (int fid = 0; ok && fid < EROD_N_FIELDS; ++fid) {
    const uint32_t nc = static_cast<uint32_t>(field_info(fid).ncomp);
    const uint64_t cnt = b->field_slots(fid);
    buf.resize(nc * cnt);
    if (int rc = erod_download_field(b, fid, buf.data())) return rc;
    ok = put(f, uint32_t(fid)) && put(f, nc) && put(f, cnt) && put_n(f, buf.data(), buf.size());
  }
  ok = ok && put_n(f, kEnd, 8);
  if (!ok) return io_error(std::string("erod_save: write failed on ") + path);
  return EROD_OK;
}

int erod_load(const char* path, erod_batch** out) {
  if (!path || !out) return io_error("erod_load: bad argument");
  *out = nullptr;
  File F;
  F.f = std::fopen(path, "rb");
  if (!F.f) return io_error(std::string("erod_load: cannot open ") + path);
  std::FILE* f = F.f;
  char magic[8];
  uint32_t version = 0, damped = 0;
  uint64_t n_rods = 0, E = 0;
  if (!get_n(f, magic, 8) || std::memcmp(magic, kMagic, 8) != 0 || !get(f, version) || version < 1 ||
      version > kVersion || !get(f, damped) || !get(f, n_rods) || !get(f, E) || n_rods == 0 || n_rods > UINT32_MAX)
    return io_error(std::string("erod_load: not an erod checkpoint (version 1 to 3): ") + path);
  // every rod takes at least its header entries and its record: refuse a count
  // the file cannot hold before sizing anything by it
  constexpr uint64_t kRodBytes = 8 + 8 + 4 + 8 + 3 * 8 + 4 + 15 * 8 + 4 + 2 * 8;
  const long here = std::ftell(f);
  if (here < 0 || std::fseek(f, 0, SEEK_END) != 0) return io_error("erod_load: cannot size the file");
  const long size = std::ftell(f);
  if (size < here || std::fseek(f, here, SEEK_SET) != 0) return io_error("erod_load: cannot size the file");
  if (n_rods > uint64_t(size - here) / kRodBytes) return io_error("erod_load: rod count exceeds the file size");
  std::vector<uint64_t> n_elems(n_rods), n_active(n_rods);
  std::vector<uint32_t> steps_left(n_rods);
  std::vector<double> t(n_rods);
  uint8_t budgets = 0;
  bool ok = get_n(f, n_elems.data(), n_rods) && get_n(f, n_active.data(), n_rods) && get_n(f, steps_left.data(), n_rods) &&
            get(f, budgets) && get_n(f, t.data(), n_rods);
  if (!ok) return io_error("erod_load: truncated header");
  erod_batch* b = nullptr;
  if (int rc = erod_batch_create(n_elems.data(), n_rods, damped ? 1 : 0, &b)) return rc;
  std::unique_ptr<erod_batch, int (*)(erod_batch*)> guard(b, erod_batch_destroy);
  if (b->E != E) return io_error("erod_load: element slots do not match the rod lengths");
  for (uint64_t r = 0; ok && r < n_rods; ++r) {
    RodInfo& ri = b->rods[r];
    int32_t ramp = 0;
    ok = get(f, ri.dt) && get(f, ri.gamma_f) && get(f, ri.gamma_t) && get(f, ri.flags) && get_n(f, ri.clamp[0].x, 3) &&
         get_n(f, ri.clamp[0].q, 9) && get_n(f, ri.force, 3) && get(f, ramp) && get(f, ri.ramp_start) && get(f, ri.ramp_interval);
    if (version >= 3) {
      ok = ok && get_n(f, ri.clamp_idx, 2) && get_n(f, ri.clamp[1].x, 3) && get_n(f, ri.clamp[1].q, 9);
    } else {
      ri.clamp_idx[0] = ri.clamp_idx[1] = 0;
      ri.flags &= ~clamp_bit(1);
    }
    ri.ramp = ramp;
    if (n_active[r] > n_elems[r] || n_active[r] == 1) ok = false;
    ri.n = static_cast<uint32_t>(n_active[r]);
    for (int k = 0; k < 2; ++k)
      if ((ri.flags & clamp_bit(k)) && ri.clamp_idx[k] > ri.n) ok = false;
    b->has_clamps |= (ri.flags & kClampBits) != 0;
    b->has_forces |= (ri.flags & 2u) != 0;
  }
  Env& e = b->env;
  int32_t grav = 0, cont = 0;
  uint32_t nf = 0;
  ok = ok && get(f, grav) && get_n(f, e.g, 3) && get(f, cont) && get(f, e.k) && get(f, e.c) && get(f, e.mu) && get(f, e.ct) &&
       get(f, e.z0) && get(f, e.offset) && (version < 2 || get_n(f, e.n, 3)) && get(f, nf) && nf == EROD_N_FIELDS;
  e.gravity = grav;
  e.contact = cont;
  if (!ok) return io_error("erod_load: bad rod records or environment");
  std::vector<double> buf;
  for (uint32_t i = 0; i < nf; ++i) {
    uint32_t fid = 0, nc = 0;
    uint64_t cnt = 0;
    if (!get(f, fid) || !get(f, nc) || !get(f, cnt) || fid >= EROD_N_FIELDS || nc != uint32_t(field_info(fid).ncomp) ||
        cnt != b->field_slots(fid))
      return io_error("erod_load: field table does not match");
    buf.resize(nc * cnt);
    if (!get_n(f, buf.data(), buf.size())) return io_error("erod_load: truncated field data");
    if (int rc = erod_upload_field(b, fid, buf.data())) return rc;
  }
  if (!get_n(f, magic, 8) || std::memcmp(magic, kEnd, 8) != 0) return io_error("erod_load: missing end marker");
  b->n_active = n_active;
  b->all_full = b->n_active == b->n_elems;
  b->steps_left = steps_left;
  if (budgets) {
    b->budgets = true;
    EROD_HIP(hipMalloc(&b->d_steps_left, sizeof(uint32_t) * n_rods));
    b->budgets_dirty = true;
  }
  b->rods_dirty = true;
  if (int rc = erod_set_rod_time(b, t.data())) return rc;
  *out = guard.release();
  return EROD_OK;
}

}  // extern "C"
