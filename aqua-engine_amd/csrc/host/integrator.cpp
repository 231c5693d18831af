// Integrator JSON (scenes/integrator.json:2-8): an internally tagged object
//   {"type": "nrc", "spp": 4, "batch_size": 512, "max_depth": 5,
//    "training_iters": 2048, "learning_rate": 0.001, "visualize_cache": false}
// plus this build's "pt" type and two optional keys, "seed" and "rr_depth".
// Unknown keys are ignored, as serde does by default.
#include <cmath>
#include <fstream>
#include <sstream>

#include "internal.h"
#include "json.h"

namespace aqua {

void integrator_default(aqua_integrator_cfg* c) {
  *c = aqua_integrator_cfg{};
  c->type = AQUA_INTEGRATOR_PT;
  c->spp = 4;              // integrator.json:3
  c->max_depth = 5;        // integrator.json:5
  c->batch_size = 512;     // integrator.json:4
  c->training_iters = 2048;// integrator.json:6
  c->learning_rate = 1e-3f;// integrator.json:7
  c->visualize_cache = 0;  // integrator.json:8
  c->seed = 0;
  c->rr_depth = 3;
  c->nrc_precision = AQUA_NRC_F16;
  c->nrc_train_paths = 1u << 16;
  c->nrc_query_vertex = 2;
  c->sampler = AQUA_SAMPLER_INDEPENDENT;
}

int integrator_load(const char* path, aqua_integrator_cfg* out) {
  if (!path || !out) return fail(AQUA_ERR_INVALID_ARG, "aqua_integrator_load: null argument");
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(AQUA_ERR_IO, "cannot open integrator '%s'", path);
  std::ostringstream ss;
  ss << f.rdbuf();
  Json j;
  std::string err;
  if (!json_parse(ss.str(), &j, &err)) return fail(AQUA_ERR_PARSE, "%s: %s", path, err.c_str());
  if (j.kind != Json::Object) return fail(AQUA_ERR_PARSE, "%s: top level must be an object", path);
  aqua_integrator_cfg c;
  integrator_default(&c);
  const Json* t = j.get("type");
  if (!t || t->kind != Json::String) return fail(AQUA_ERR_PARSE, "%s: missing string 'type'", path);
  if (t->str == "pt")
    c.type = AQUA_INTEGRATOR_PT;
  else if (t->str == "nrc")
    c.type = AQUA_INTEGRATOR_NRC;
  else
    return fail(AQUA_ERR_PARSE, "%s: unknown integrator type '%s'", path, t->str.c_str());
  struct U {
    const char* key;
    uint32_t* dst;
    double lo, hi;
  } ints[] = {{"spp", &c.spp, 1, 1 << 20},
              {"max_depth", &c.max_depth, 0, 64},
              {"batch_size", &c.batch_size, 1, 1 << 30},
              {"training_iters", &c.training_iters, 0, 1 << 30},
              {"seed", &c.seed, 0, 4294967295.0},
              {"rr_depth", &c.rr_depth, 0, 1 << 20},
              {"nrc_train_paths", &c.nrc_train_paths, 1, 1 << 22},
              {"nrc_query_vertex", &c.nrc_query_vertex, 1, 64}};
  for (const U& u : ints) {
    const Json* v = j.get(u.key);
    if (!v) continue;
    if (!v->is_num() || v->num != std::floor(v->num) || v->num < u.lo || v->num > u.hi)
      return fail(AQUA_ERR_PARSE, "%s: '%s' must be an integer in [%g, %g]", path, u.key, u.lo, u.hi);
    *u.dst = uint32_t(v->num);
  }
  if (const Json* v = j.get("learning_rate")) {
    if (!v->is_num() || !(v->num > 0.0) || !std::isfinite(v->num))
      return fail(AQUA_ERR_PARSE, "%s: 'learning_rate' must be a positive number", path);
    c.learning_rate = float(v->num);
  }
  if (const Json* v = j.get("nrc_precision")) {
    if (v->kind != Json::String) return fail(AQUA_ERR_PARSE, "%s: 'nrc_precision' must be a string", path);
    if (v->str == "f16")
      c.nrc_precision = AQUA_NRC_F16;
    else if (v->str == "bf16")
      c.nrc_precision = AQUA_NRC_BF16;
    else if (v->str == "fp8")
      c.nrc_precision = AQUA_NRC_FP8;
    else
      return fail(AQUA_ERR_PARSE, "%s: unknown nrc_precision '%s'", path, v->str.c_str());
  }
  if (const Json* v = j.get("sampler")) {
    if (v->kind != Json::String) return fail(AQUA_ERR_PARSE, "%s: 'sampler' must be a string", path);
    if (v->str == "independent")
      c.sampler = AQUA_SAMPLER_INDEPENDENT;
    else if (v->str == "stratified")
      c.sampler = AQUA_SAMPLER_STRATIFIED;
    else
      return fail(AQUA_ERR_PARSE, "%s: unknown sampler '%s'", path, v->str.c_str());
  }
  if (const Json* v = j.get("visualize_cache")) {
    if (v->kind != Json::Bool) return fail(AQUA_ERR_PARSE, "%s: 'visualize_cache' must be a bool", path);
    c.visualize_cache = v->b ? 1u : 0u;
  }
  *out = c;
  return AQUA_OK;
}

}  // namespace aqua
