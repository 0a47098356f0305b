// tuning_table.hpp -- launch settings of the tiled contractions per
// molecule-size bucket, written by occ_amd/autotune.py from timed builds on
// one MI355X (python -m occ_amd.autotune).  An entry applies to contexts
// whose largest molecule has at most max_nbf basis functions.
#pragma once

namespace occxc {

struct TuneEntry {
  int max_nbf;     // bucket: largest molecule's basis functions
  int tile_ao;     // largest AO block per work item
  int persistent;  // bit 0 persistent density, bit 1 persistent Vxc
};

static const TuneEntry kTuneTable[] = {
    {64, 96, 0},  // 48 molecules of at most 64 AOs (batch, 50x194): 3.582 ms (default 128/0: 3.591 ms)
    {256, 128, 0},  // 32 molecules of 65-256 AOs (batch, 50x194): 16.181 ms (default 128/0: 16.181 ms)
    {1024, 128, 0},  // (H2O)32 / def2-SVP, 75x302: 9.265 ms (default 128/0: 9.265 ms)
    {1 << 30, 128, 0},  // (H2O)128 / def2-SVP, 75x302: 73.046 ms (default 128/0: 73.046 ms)
};

} // namespace occxc
