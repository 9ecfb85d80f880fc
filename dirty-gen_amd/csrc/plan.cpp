// Host-side work-box and substep planning (see plan.hpp).
#include "plan.hpp"

#include <math.h>

#include <algorithm>
#include <numeric>

namespace silt {

namespace {

struct Run {
  int a, b;
};

int uf_root(std::vector<uint32_t>& parent, uint32_t x) {
  while (parent[x] != x) {
    uint32_t up = parent[parent[x]];
    parent[x] = up;
    x = up;
  }
  return (int)x;
}

bool touch(const Box& a, const Box& b) {
  // boxes_touch (src/fluid_core.rs:1449-1456): within one cell, saturating at 0.
  int ai0 = a.i0 > 0 ? a.i0 - 1 : 0, bi0 = b.i0 > 0 ? b.i0 - 1 : 0;
  int aj0 = a.j0 > 0 ? a.j0 - 1 : 0, bj0 = b.j0 > 0 ? b.j0 - 1 : 0;
  return ai0 <= b.i1 + 1 && bi0 <= a.i1 + 1 && aj0 <= b.j1 + 1 && bj0 <= a.j1 + 1;
}

// merge_boxes (src/fluid_core.rs:1458-1486).
void merge(std::vector<Box>& boxes) {
  for (;;) {
    bool did = false;
    for (size_t i = 0; i < boxes.size(); i++) {
      size_t j = i + 1;
      while (j < boxes.size()) {
        if (touch(boxes[i], boxes[j])) {
          Box a = boxes[i], b = boxes[j];
          boxes[i] = Box{std::min(a.i0, b.i0), std::max(a.i1, b.i1), std::min(a.j0, b.j0), std::max(a.j1, b.j1)};
          boxes[j] = boxes.back();
          boxes.pop_back();
          did = true;
        } else {
          j++;
        }
      }
    }
    if (!did) break;
  }
}

}  // namespace

std::vector<Box> sheet_work_boxes(const float* h, int res, float eps, bool allow_split) {
  int i0 = res, i1 = 0, j0 = res, j1 = 0;
  long wet = 0;
  for (int j = 0; j < res; j++) {
    const float* row = h + (size_t)j * res;
    int first = -1, last = -1, c = 0;
    for (int i = 0; i < res; i++)
      if (row[i] > eps) {
        if (first < 0) first = i;
        last = i;
        c++;
      }
    if (!c) continue;
    wet += c;
    i0 = std::min(i0, first);
    i1 = std::max(i1, last);
    j0 = std::min(j0, j);
    j1 = std::max(j1, j);
  }
  if (!wet) return {};
  Box halo3{std::max(i0 - 3, 0), std::min(i1 + 3, res - 1), std::max(j0 - 3, 0), std::min(j1 + 3, res - 1)};
  bool solid = wet == (long)(i1 - i0 + 1) * (long)(j1 - j0 + 1);
  if (!allow_split || solid) return {halo3};
  // 4-connected wet bodies from row runs, union by the older run.
  std::vector<Run> runs;
  std::vector<uint32_t> parent;
  std::vector<Box> bbox;
  std::vector<size_t> first;
  size_t prev0 = 0, prev1 = 0;
  for (int j = j0; j <= j1; j++) {
    const float* row = h + (size_t)j * res;
    size_t start = runs.size();
    int i = i0;
    while (i <= i1) {
      if (!(row[i] > eps)) {
        i++;
        continue;
      }
      int a = i;
      while (i <= i1 && row[i] > eps) i++;
      int b = i - 1;
      uint32_t id = (uint32_t)runs.size();
      runs.push_back({a, b});
      parent.push_back(id);
      bbox.push_back(Box{a, b, j, j});
      first.push_back((size_t)j * res + a);
      for (size_t k = prev0; k < prev1; k++) {
        if (runs[k].a <= b && a <= runs[k].b) {
          uint32_t ra = (uint32_t)uf_root(parent, (uint32_t)k), rb = (uint32_t)uf_root(parent, id);
          if (ra != rb) {
            uint32_t keep = std::min(ra, rb), gone = std::max(ra, rb);
            parent[gone] = keep;
            Box g = bbox[gone];
            Box& kb = bbox[keep];
            kb.i0 = std::min(kb.i0, g.i0);
            kb.i1 = std::max(kb.i1, g.i1);
            kb.j0 = std::min(kb.j0, g.j0);
            kb.j1 = std::max(kb.j1, g.j1);
            first[keep] = std::min(first[keep], first[gone]);
          }
        }
      }
    }
    prev0 = start;
    prev1 = runs.size();
  }
  std::vector<std::pair<size_t, Box>> bodies;
  for (size_t k = 0; k < runs.size(); k++)
    if (parent[k] == (uint32_t)k) bodies.push_back({first[k], bbox[k]});
  std::sort(bodies.begin(), bodies.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
  std::vector<Box> bb;
  for (auto& bd : bodies) bb.push_back(bd.second);
  return boxes_from_bodies(std::move(bb), res, halo3);
}

// The merged set does not depend on the bodies' order (merging only grows
// boxes, so bodies whose boxes touch at any point end in one box whatever the
// order); only the list order does, and boxes that do not touch are stepped
// independently.
std::vector<Box> boxes_from_bodies(std::vector<Box> bodies, int res, const Box& halo3) {
  std::vector<Box> boxes;
  for (const Box& b : bodies)
    boxes.push_back(Box{std::max(b.i0 - 6, 0), std::min(b.i1 + 6, res - 1), std::max(b.j0 - 6, 0),
                        std::min(b.j1 + 6, res - 1)});
  merge(boxes);
  if (boxes.size() == 1 || boxes.size() > 8) return {halo3};
  return boxes;
}

std::vector<Box> awake_work_boxes(const float* h, const float* u, const float* v, const float* bed, int res,
                                  float cell, float eps) {
  const int T = 8;
  int tiles = (res + T - 1) / T;
  std::vector<uint8_t> hot((size_t)tiles * tiles, 0);
  for (int tj = 0; tj < tiles; tj++)
    for (int ti = 0; ti < tiles; ti++) {
      int ia = ti * T, ja = tj * T, ib = std::min(ia + T, res), jb = std::min(ja + T, res);
      bool awake = false;
      for (int j = ja; j < jb && !awake; j++)
        for (int i = ia; i < ib; i++) {
          size_t idx = (size_t)j * res + i;
          if (h[idx] <= eps) continue;
          if (hypotf(u[idx], v[idx]) > 0.05f) {
            awake = true;
            break;
          }
          float eta = bed[idx] + h[idx];
          size_t ip = i + 1 < res ? idx + 1 : idx;
          size_t jp = j + 1 < res ? idx + res : idx;
          float dex = fabsf(bed[ip] + h[ip] - eta);
          float dez = fabsf(bed[jp] + h[jp] - eta);
          if (fmaxf(dex, dez) > 0.02f * cell) {
            awake = true;
            break;
          }
        }
      if (awake) hot[(size_t)tj * tiles + ti] = 1;
    }
  return awake_boxes_from_hot(hot, res);
}

std::vector<Box> awake_boxes_from_hot(const std::vector<uint8_t>& hot, int res) {
  const int T = 8;
  int tiles = (res + T - 1) / T;
  std::vector<uint8_t> grown((size_t)tiles * tiles, 0), seen((size_t)tiles * tiles, 0);
  for (int tj = 0; tj < tiles; tj++)
    for (int ti = 0; ti < tiles; ti++) {
      if (!hot[(size_t)tj * tiles + ti]) continue;
      for (int dj = -1; dj <= 1; dj++)
        for (int di = -1; di <= 1; di++) {
          int ni = ti + di, nj = tj + dj;
          if (ni >= 0 && nj >= 0 && ni < tiles && nj < tiles) grown[(size_t)nj * tiles + ni] = 1;
        }
    }
  std::vector<Box> boxes;
  std::vector<int> stack;
  for (int s = 0; s < tiles * tiles; s++) {
    if (!grown[s] || seen[s]) continue;
    seen[s] = 1;
    stack.assign(1, s);
    int ti0 = s % tiles, ti1 = ti0, tj0 = s / tiles, tj1 = tj0;
    while (!stack.empty()) {
      int cur = stack.back();
      stack.pop_back();
      int tj = cur / tiles, ti = cur - tj * tiles;
      ti0 = std::min(ti0, ti);
      ti1 = std::max(ti1, ti);
      tj0 = std::min(tj0, tj);
      tj1 = std::max(tj1, tj);
      int nb[4], nn = 0;
      if (ti > 0) nb[nn++] = cur - 1;
      if (ti + 1 < tiles) nb[nn++] = cur + 1;
      if (tj > 0) nb[nn++] = cur - tiles;
      if (tj + 1 < tiles) nb[nn++] = cur + tiles;
      for (int q = 0; q < nn; q++)
        if (grown[nb[q]] && !seen[nb[q]]) {
          seen[nb[q]] = 1;
          stack.push_back(nb[q]);
        }
    }
    boxes.push_back(Box{ti0 * T, std::min((ti1 + 1) * T, res) - 1, tj0 * T, std::min((tj1 + 1) * T, res) - 1});
  }
  return boxes;
}

bool reference_uses_jacobi(const Box& b, float cell, int wind, bool row_split) {
  bool fine = !wind && cell <= kFineCell + 1.0e-6f;
  bool big = (b.j1 - b.j0 >= 47) && (b.i1 - b.i0 >= 31) && row_split;
  return fine || big;
}

StepPlan plan_step(const float* h, const float* u, const float* v, const float* bed, int res, float cell, float dt,
                   int wind, float wx, float wz, float eps, float sub_cap, int pace, bool walled) {
  StepPlan plan;
  if (!(dt > 0.0f) || res <= 0) return plan;
  // wet_sum (src/fluid_core.rs:1305-1319), eight lanes then the remainder.
  size_t n = (size_t)res * res, full = n / 8 * 8;
  float lanes[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (size_t c = 0; c < full; c += 8)
    for (int k = 0; k < 8; k++) lanes[k] += h[c + k] > eps ? h[c + k] : 0.0f;
  float sum = -0.0f;
  for (int k = 0; k < 8; k++) sum += lanes[k];
  for (size_t c = full; c < n; c++) sum += h[c] > eps ? h[c] : 0.0f;
  if (sum <= eps) return plan;
  // A wall keeps the wet-set boxes and the substep loop on every grid
  // (step_sheet_scratch :1694-1696) and Courant 0.42 for wind too (:1735-1738).
  bool open_fine = !walled && !wind && cell <= kFineCell + 1.0e-6f;
  plan.fine_open = open_fine && pace == 0;
  bool whole_wet = open_fine && (pace == 2 || pace == 3);
  bool awake = open_fine && !whole_wet && !plan.fine_open;
  plan.boxes = (plan.fine_open || awake) ? awake_work_boxes(h, u, v, bed, res, cell, eps)
                                         : sheet_work_boxes(h, res, eps, true);
  if (plan.boxes.empty()) return plan;
  float cap = isfinite(sub_cap) ? clampf(sub_cap, 1.0f, 8.0f) : 8.0f;
  if (plan.fine_open) {
    plan.nsub = 1;
  } else {
    float max_s = 0.15f;
    if (wind) max_s = fmaxf(max_s, hypotf(wx, wz));
    float max_s2 = 0.0f;
    for (const Box& b : plan.boxes)
      for (int j = b.j0; j <= b.j1; j++)
        for (int i = b.i0; i <= b.i1; i++) {
          size_t idx = (size_t)j * res + i;
          float s2 = u[idx] * u[idx] + v[idx] * v[idx];
          if (s2 > max_s2) max_s2 = s2;
        }
    max_s = fmaxf(max_s, sqrtf(max_s2));
    float courant = wind && !walled ? 0.85f : 0.42f;
    float q = ceilf((dt * max_s) / (courant * fmaxf(cell, 1.0e-3f)));
    plan.nsub = (int)clampf(q, 1.0f, cap);
  }
  plan.pace3 = pace == 3 && whole_wet && plan.nsub > 1;
  return plan;
}

}  // namespace silt
