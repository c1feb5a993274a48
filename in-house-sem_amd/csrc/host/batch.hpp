// Many independent models as one setup: the disjoint union of their
// domains, blocks, points, boundary data, coupling, sources and receivers.
//
// The union is what the batched backend steps: one element-force launch per
// element group and one tail per domain cover every model at once, instead
// of one set of launches per model.  Models keep their own time step and
// step count; every per-point computation only touches its own model's
// data, and each point's assembly sum keeps its model's slot order, so a
// model's results in the union are bitwise its results alone.
#pragma once

#include <memory>
#include <vector>

#include "model.hpp"

namespace insem_amd {

struct BatchImage {
  std::unique_ptr<SetupImage> image;            // the union
  std::vector<std::vector<int32_t>> point_model;  // per domain: model of each global point
  std::vector<int32_t> receiver_model;           // model of each union receiver
  std::vector<int32_t> source_model;             // model of each union source
  std::vector<int32_t> receiver_offset;          // first union receiver of each model
  int n_models = 0;
};

// All models must share the polynomial order, the coordinate system and
// azimuthal order, the media present and record_every.
BatchImage union_of(const std::vector<const SetupImage*>& models);

}  // namespace insem_amd
