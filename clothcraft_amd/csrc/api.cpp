// api.cpp — solver-side entry points of the C ABI (include/clothvbd.h).
#include <cstdio>
#include <cstring>

#include "solver.hpp"

thread_local std::string cv_create_error;

namespace {
template <class F>
cv_status guard(std::string* err, F&& f) {
  try {
    f();
    return CV_OK;
  } catch (const CvError& e) {
    if (err) *err = e.message;
    return e.status;
  } catch (const std::bad_alloc&) {
    if (err) *err = "out of host memory";
    return CV_ERR_CAPACITY;
  } catch (const std::exception& e) {
    if (err) *err = e.what();
    return CV_ERR_INVALID_ARGUMENT;
  }
}
}  // namespace

// GPU handles keep their own last error; see vbd_device.hip.
const std::string* cv_gpu_error_string(const cv_gpu* g);

extern "C" {

cv_status cv_solver_create(const cv_garment* garment, const cv_surface* avatar, const cv_state* state,
                           cv_solver** out) {
  if (!out) return CV_ERR_INVALID_ARGUMENT;
  cv_solver* s = new cv_solver();
  cv_status st = guard(&cv_create_error, [&] { build_solver(*s, garment, avatar, state); });
  if (st != CV_OK) {
    delete s;
    *out = nullptr;
    return st;
  }
  *out = s;
  return CV_OK;
}

void cv_solver_destroy(cv_solver* s) { delete s; }

cv_status cv_solver_record_prediction(cv_solver* s, const double* old_velocity, const cv_state* st,
                                      const cv_settings* set) {
  if (!s || !old_velocity || !st || !set) return CV_ERR_INVALID_ARGUMENT;
  return guard(&s->last_error, [&] {
    validate_state(*s, st);
    size_t n = s->n;
    s->acc.resize(3 * n);
    for (size_t i = 0; i < 3 * n; ++i)
      s->acc[i] = (st->v[i] - old_velocity[i]) / set->dt + set->damping * st->v[i];
    s->has_last = true;
    s->last_x.assign(st->x, st->x + 3 * n);
    s->last_v.assign(st->v, st->v + 3 * n);
    s->last_time = st->time;
    for (int k = 0; k < 3; ++k) s->last_gravity[k] = set->gravity[k];
    s->last_damping = set->damping;
    static uint64_t host_version = 1;
    s->acc_version = ++host_version;
  });
}

cv_status cv_solver_colors(const cv_solver* s, uint32_t* color_of_vertex, uint32_t n, uint32_t* n_colors) {
  if (!s || !color_of_vertex || n != s->n) return CV_ERR_INVALID_ARGUMENT;
  std::vector<uint32_t> col, order, off;
  uint32_t k = greedy_colors(*s, nullptr, nullptr, nullptr, col, order, off);
  std::memcpy(color_of_vertex, col.data(), n * sizeof(uint32_t));
  if (n_colors) *n_colors = k;
  return CV_OK;
}

size_t cv_last_error(const cv_gpu* gpu, char* buf, size_t n) {
  const std::string* msg = &cv_create_error;
  if (gpu) msg = cv_gpu_error_string(gpu);
  if (buf && n) std::snprintf(buf, n, "%s", msg->c_str());
  return msg->size();
}

size_t cv_solver_last_error(const cv_solver* s, char* buf, size_t n) {
  const std::string& msg = s ? s->last_error : cv_create_error;
  if (buf && n) std::snprintf(buf, n, "%s", msg.c_str());
  return msg.size();
}

const char* cv_version(void) { return "clothvbd 0.1.0 (gfx950)"; }

}  // extern "C"
