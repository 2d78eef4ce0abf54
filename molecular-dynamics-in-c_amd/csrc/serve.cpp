// serve.cpp -- placeholder replaced by the serving loop
#include "mdgpu.h"
extern "C" int mdg_serve(const mdg_job*, int, const mdg_serve_config*, mdg_job_result*,
                         mdg_serve_stats*) {
  return MDG_E_ARG;
}
