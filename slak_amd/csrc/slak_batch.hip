// Batched subdivision: thousands of meshes of uneven size and depth refined
// together, one launch per kernel per level over the union of the meshes
// still active at that level, the whole chain replayed as one hipGraph.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <functional>
#include <memory>

#include "slak_c.h"
#include "slak_plan.hpp"

struct slak_batch {
  std::unique_ptr<slak::Plan> plan;
  int scheme = 0;
  uint64_t out_faces = 0, out_verts = 0;
};

namespace slak {
int guarded_call(const std::function<void()>& fn);
}

extern "C" {

int slak_batch_create(int32_t scheme, int32_t nmeshes, const slak_mesh* meshes, const int32_t* target_levels,
                      slak_batch** batch, void* stream) {
  return slak::guarded_call([&] {
    if (!batch || (nmeshes > 0 && (!meshes || !target_levels)) || nmeshes < 0)
      throw slak::Error(SLAK_ERR_INVALID_ARGUMENT, "null argument");
    if (scheme != SLAK_SCHEME_LAK && scheme != SLAK_SCHEME_SLAK)
      throw slak::Error(SLAK_ERR_INVALID_ARGUMENT, "unknown scheme");
    *batch = nullptr;
    std::vector<slak::MeshJob> jobs(nmeshes);
    for (int i = 0; i < nmeshes; ++i) {
      jobs[i].in = meshes[i];
      jobs[i].target = target_levels[i];
    }
    auto b = std::make_unique<slak_batch>();
    b->scheme = scheme;
    b->plan.reset(new slak::Plan(scheme, std::move(jobs), slak::OutputMode::kPooled, static_cast<hipStream_t>(stream)));
    b->plan->prepare();
    for (auto& J : b->plan->jobs())
      if (J.levels > 0) {
        b->out_faces += (uint64_t)J.nf[J.levels];
        b->out_verts += (uint64_t)J.nv[J.levels];
      }
    *batch = b.release();
  });
}

int slak_batch_run(slak_batch* batch, int32_t use_graph) {
  return slak::guarded_call([&] {
    if (!batch || !batch->plan) throw slak::Error(SLAK_ERR_INVALID_ARGUMENT, "null batch");
    if (use_graph)
      batch->plan->run_graph();
    else
      batch->plan->launch(true);
    batch->plan->finish();
  });
}

int slak_batch_result(slak_batch* batch, int32_t index, slak_mesh* rmesh) {
  return slak::guarded_call([&] {
    if (!batch || !batch->plan || !rmesh) throw slak::Error(SLAK_ERR_INVALID_ARGUMENT, "null argument");
    auto& jobs = batch->plan->jobs();
    if (index < 0 || index >= (int)jobs.size()) throw slak::Error(SLAK_ERR_INVALID_ARGUMENT, "index out of range");
    const slak::MeshJob& J = jobs[index];
    if (J.levels <= 0) {  // nothing refined: the output stays untouched
      return;
    }
    const int L = J.levels;
    rmesh->type = SLAK_QUAD;
    rmesh->nfaces = (uint64_t)J.nf[L];
    rmesh->nnz = (uint64_t)J.nnz[L];
    rmesh->nverts = (uint64_t)J.nv[L];
    rmesh->ids = J.out_ids;
    rmesh->verts = reinterpret_cast<float*>(J.out_verts);
    if (batch->scheme == SLAK_SCHEME_LAK) {
      rmesh->is_reduced = 0;
      rmesh->max_face_order = 4;
      rmesh->ptr = J.out_ptr;
      rmesh->vals = J.out_vals;
    } else {
      rmesh->is_reduced = 1;
      rmesh->ptr = nullptr;
      rmesh->vals = nullptr;
    }
  });
}

int slak_batch_info_get(slak_batch* batch, slak_batch_info* info) {
  return slak::guarded_call([&] {
    if (!batch || !batch->plan || !info) throw slak::Error(SLAK_ERR_INVALID_ARGUMENT, "null argument");
    std::memset(info, 0, sizeof(*info));
    info->nmeshes = (int32_t)batch->plan->jobs().size();
    info->max_levels = batch->plan->max_levels();
    info->host_syncs = batch->plan->syncs();
    info->device_bytes = (uint64_t)batch->plan->bytes();
    info->out_faces = batch->out_faces;
    info->out_verts = batch->out_verts;
  });
}

int slak_batch_destroy(slak_batch* batch) {
  return slak::guarded_call([&] { delete batch; });
}

}  // extern "C"
