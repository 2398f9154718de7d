// Drop-in for hcorepp::operators::TilePacker<T> (reference src/operators/TilePacker.cpp:1-28):
// a tile <-> (TileMetadata, data pointer) wire format for task runtimes.  PackTile returns a
// non-owning tile over the caller's buffer; the caller deletes the returned objects.
#pragma once

#include <hcorepp/operators/concrete/Compressed.hpp>
#include <hcorepp/operators/concrete/Dense.hpp>

namespace hcorepp {
namespace operators {

template <typename T>
class TilePacker {
public:
    static std::pair<TileMetadata *, T *> UnPackTile(const Tile<T> &aTile, const kernels::RunContext &aContext) {
        return aTile.UnPackTile(aContext);
    }
    static Tile<T> *PackTile(const TileMetadata &aMetadata, T *apDataArray, const kernels::RunContext &) {
        if (aMetadata.mType == DENSE) return new DenseTile<T>(aMetadata, apDataArray);
        return new CompressedTile<T>(aMetadata, apDataArray);
    }
};

}  // namespace operators
}  // namespace hcorepp
