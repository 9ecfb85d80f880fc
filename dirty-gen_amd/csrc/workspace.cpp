#include "workspace.hpp"

#include <map>
#include <tuple>
#include <memory>
#include <utility>

namespace silt {

Workspace* workspace_for(int res, int slot) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int>, std::unique_ptr<Workspace>> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  auto key = std::make_tuple(dev, res, slot);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second.get();
  auto ws = std::make_unique<Workspace>();
  ws->device = dev;
  ws->res = res;
  ws->n = (size_t)res * res;
  if (hipMalloc(&ws->fbuf, (size_t)kWorkspacePlanes * ws->n * sizeof(float)) != hipSuccess) return nullptr;
  if (hipMalloc(&ws->gate, 2 * ws->n) != hipSuccess) return nullptr;
  if (hipMemset(ws->fbuf, 0, (size_t)kWorkspacePlanes * ws->n * sizeof(float)) != hipSuccess) return nullptr;
  if (hipStreamCreateWithFlags(&ws->stream, hipStreamNonBlocking) != hipSuccess) return nullptr;
  Workspace* out = ws.get();
  cache.emplace(key, std::move(ws));
  return out;
}

}  // namespace silt
