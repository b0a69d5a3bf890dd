// Test harness: a recording stand-in for the C ABI calls erod_elastica.hpp
// makes (include/erod.h), so the stepper surface can be checked on a host
// with no GPU.  Each call appends "name arg..." to a log; erod_fill_field on
// field -7 fails, to check that failures reach the caller.
#include <cstdio>
#include <string>

#include "../include/erod.h"

static std::string g_log;

static int rec(const char* name, const erod_batch* b, const char* fmt, double x, double y = 0.0) {
  char buf[256];
  std::snprintf(buf, sizeof(buf), fmt, x, y);
  g_log += std::string(name) + " " + std::to_string(reinterpret_cast<unsigned long>(b)) + " " + buf + "\n";
  return EROD_OK;
}

extern "C" {
const char* erod_last_error(void) { return "stub failure"; }
int erod_kinematic_update_by(erod_batch* b, double c) { return rec("kinematic_update_by", b, "%.17g", c); }
int erod_dynamic_update(erod_batch* b, double dt) { return rec("dynamic_update", b, "%.17g", dt); }
int erod_compute_internal_dynamics(erod_batch* b, double t) { return rec("compute_internal_dynamics", b, "%.17g", t); }
int erod_apply_post_kinematic(erod_batch* b, double t) { return rec("apply_post_kinematic", b, "%.17g", t); }
int erod_apply_pre_dynamic(erod_batch* b, double t) { return rec("apply_pre_dynamic", b, "%.17g", t); }
int erod_apply_post_dynamic(erod_batch* b, double t) { return rec("apply_post_dynamic", b, "%.17g", t); }
int erod_fill_field(erod_batch* b, int field, double v) {
  if (field == -7) return EROD_ERR_INVALID;
  return rec("fill_field", b, "%.0f %.17g", double(field), v);
}
const char* stub_log(void) { return g_log.c_str(); }
void stub_clear(void) { g_log.clear(); }
}
