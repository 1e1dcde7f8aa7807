// Host construction of the environment-map sampling tables.
#pragma once
#include <stdint.h>

#include <vector>

namespace rgpu {

struct AliasBucket {
    float height;
    int32_t alias;  // -1: no alias
};

struct EnvFragment {
    float beg_y, end_y, beg_phi, end_phi;
};

// Tables for one RGB image of width x height texels (row 0 at the zenith).
// table 0 is the uniform-weight table, tables 1..6 the six-axis tables (+X,-X,+Y,-Y,+Z,-Z).
struct EnvmapTables {
    uint32_t width = 0, height = 0;
    std::vector<float> texels;  // RGBA per texel (A = 0)
    std::vector<EnvFragment> fragments;
    std::vector<float> pdf[7];
    std::vector<AliasBucket> alias[7];
    bool degenerate[7] = {false, false, false, false, false, false, false};
};

// rgb: width*height*channels floats (channels 3 or 4, row-major from the top row).
void build_envmap_tables(const float* rgb, uint32_t channels, uint32_t width, uint32_t height, EnvmapTables* out);

}  // namespace rgpu
