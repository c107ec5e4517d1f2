#include "fe_internal.hpp"
extern "C" {
void a2c_lmvm_opts_default(a2c_lmvm_opts* o) {
  o->m = 5; o->max_it = 2000; o->max_funcs = 4000; o->gatol = 1e-8; o->grtol = 1e-8; o->gttol = 0.0;
  o->sigma_lb = 1e-6; o->mixed_switch = 1; o->mixed_gtol = 1e-5;
}
int a2c_fe_minimize(a2c_fe_ctx* c, const double*, const a2c_lmvm_opts*, double*, a2c_lmvm_report*) {
  return a2c::set_err(c, A2C_ESTATE, "not available yet");
}
int a2c_fe_chem_potential(a2c_fe_ctx* c, const double*, const double*, const double*, double, double*, int32_t*) {
  return a2c::set_err(c, A2C_ESTATE, "not available yet");
}
int a2c_fe_diffuse(a2c_fe_ctx* c, double, double, double, double, double, const double*, const int32_t*, double*) {
  return a2c::set_err(c, A2C_ESTATE, "not available yet");
}
}
