// Host tiling of the D&C (svd_tiling.cpp:67-130): the reference's blob is an Ascend
// SVDTiling (TCubeTiling + stack size + offset) followed by uint16 leaf ranges.  This
// blob keeps the two entry points and the leaf partition; the Ascend GEMM descriptor
// has no meaning on gfx950 and is replaced by a version header.  Leaf ranges are
// uint32 (the reference's uint16 limits N to 65535, svd.cpp:60-64).
//   bytes 0..63   header: magic "PACO", version (2), N, leafCount, leafOffset (=64),
//                 svdBlockNum, leafSize, reserved
//   bytes 64..    leafCount x {uint32 start_col, uint32 end_col}
// The leaf size is the library's current one (paco_leaf_size; size 3 gives the
// reference's partition): paco_svd_DC checks a blob against the partition it uses.
#include <cstdint>
#include <cstring>

#include "../../include/paco_svd.h"
#include "paco_common.hpp"

namespace {
struct TilingHeader {
    uint32_t magic, version, N, leafCount, leafOffset, svdBlockNum, leafSize;
    uint32_t reserved[9];
};
static_assert(sizeof(TilingHeader) == 64, "tiling header is 64 bytes");
} // namespace

extern "C" int getTilingSize(int N, int svdBlockNum)
{
    (void)svdBlockNum;
    if (N < 1) return (int)sizeof(TilingHeader);
    return (int)(sizeof(TilingHeader) + 8 * (size_t)paco::leaf_count(N, paco_leaf_size()));
}

extern "C" void GenerateTiling(int N, int svdBlockNum, uint8_t *tiling_host)
{
    if (!tiling_host) return;
    TilingHeader h;
    memset(&h, 0, sizeof(h));
    const int ls = paco_leaf_size();
    h.magic = 0x4F434150u; // "PACO"
    h.version = 2;
    h.N = (uint32_t)(N > 0 ? N : 0);
    h.leafCount = N > 0 ? (uint32_t)paco::leaf_count(N, ls) : 0;
    h.leafOffset = sizeof(TilingHeader);
    h.svdBlockNum = (uint32_t)svdBlockNum;
    h.leafSize = (uint32_t)ls;
    memcpy(tiling_host, &h, sizeof(h));
    uint32_t *leaves = reinterpret_cast<uint32_t *>(tiling_host + sizeof(TilingHeader));
    for (uint32_t l = 0; l < h.leafCount; ++l) {
        leaves[2 * l] = (uint32_t)paco::leaf_start(N, (int)l, ls);
        leaves[2 * l + 1] = (uint32_t)paco::leaf_start(N, (int)l + 1, ls);
    }
}
