// Worker threads shared by the batch drivers: each worker owns one context
// (HIP stream + scratch buffers) on the pool's device.
#pragma once
#include <functional>
#include <vector>

#include "ctx.hpp"

namespace cpsz {

constexpr int kMaxWorkers = 16;

class WorkerPool;
// The device's pool of kMaxWorkers workers (created on first use, kept until
// the process ends: its threads and contexts must not be torn down after the
// HIP runtime at exit).
WorkerPool& pool_for(int device);
// Runs every job on at most `limit` workers at once and returns when all have
// finished; rethrows the first failure. Calls from different host threads
// are serialized.
void pool_run(WorkerPool& pool, std::vector<std::function<void(Ctx&)>>& jobs, int limit);

}  // namespace cpsz
