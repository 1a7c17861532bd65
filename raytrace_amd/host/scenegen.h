// scenegen.h — seeded synthetic render workloads (see scenegen.cpp).
#pragma once
#include "Scene.h"
#include "../../include/rtg_host.h"

// Returns a new Scene (caller owns) or nullptr for an unknown workload.
Scene *rtg_make_config_scene(int config, const rtg_gen_opts &opts, rtg_gen_info &info);
void rtg_make_requests(uint64_t seed, int n, double rate, rtg_gen_request *out);
