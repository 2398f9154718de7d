// Drop-in for hcorepp::operators::Tile / TileMetadata / TileType
// (reference include/hcorepp/operators/interface/Tile.hpp:21-294).
#pragma once

#include <cstddef>
#include <cstring>
#include <utility>
#include <stdexcept>

#include <hcorepp/common/Definitions.hpp>
#include <hcorepp/kernels/RunContext.hpp>
#include <hcorepp_amd.h>

namespace hcorepp {
namespace operators {

enum TileType { DENSE, COMPRESSED };

struct TileMetadata {
    size_t mNumOfRows, mNumOfCols, mMatrixRank, mMaxRank, mLeadingDimension;
    blas::Layout mLayout;
    TileType mType;
    TileMetadata(size_t aRows, size_t aCols, size_t aMatrixRank, size_t aMaxRank, size_t aLeadingDim,
                 blas::Layout aLayout, TileType aType)
        : mNumOfRows(aRows), mNumOfCols(aCols), mMatrixRank(aMatrixRank), mMaxRank(aMaxRank),
          mLeadingDimension(aLeadingDim), mLayout(aLayout), mType(aType) {}
};

// Base of DenseTile / CompressedTile: owns a device buffer (unless wrapping caller memory)
// and the descriptor handed to the C-ABI.
template <typename T>
class Tile {
public:
    virtual ~Tile() {
        if (mOwns && mDesc.data) hcorepp_free(mDesc.data);
    }
    size_t GetNumOfRows() const { return (size_t) mDesc.rows; }
    size_t GetNumOfCols() const { return (size_t) mDesc.cols; }
    size_t GetLeadingDim() const { return (size_t) mDesc.ld; }
    blas::Layout GetLayout() const { return blas::Layout::ColMajor; }
    bool isDense() const { return mDesc.type == HCOREPP_DENSE; }
    bool isCompressed() const { return mDesc.type == HCOREPP_COMPRESSED; }
    int GetNumOfSubMatrices() const { return isDense() ? 1 : 2; }
    hcorepp_tile &Desc() { return mDesc; }
    const hcorepp_tile &Desc() const { return mDesc; }
    T *GetData() const { return static_cast<T *>(mDesc.data); }
    // Metadata + data pointer of the tile (reference Tile::UnPackTile, Dense.cpp:127-134,
    // Compressed.cpp:769-779): the caller owns the returned TileMetadata; the data stays the tile's.
    std::pair<TileMetadata *, T *> UnPackTile(const kernels::RunContext &) const {
        return {new TileMetadata((size_t) mDesc.rows, (size_t) mDesc.cols, (size_t) mDesc.rank,
                                 (size_t) mDesc.max_rank, (size_t) mDesc.ld, blas::Layout::ColMajor,
                                 mDesc.type == HCOREPP_DENSE ? DENSE : COMPRESSED),
                GetData()};
    }

protected:
    void Allocate(size_t aElements) {
        void *p = nullptr;
        kernels::ThrowIfError(hcorepp_malloc(aElements * sizeof(T), &p));
        kernels::ThrowIfError(hcorepp_memset(nullptr, p, 0, aElements * sizeof(T)));
        mDesc.data = p;
        mOwns = true;
    }
    hcorepp_tile mDesc{};
    bool mOwns = false;
};

}  // namespace operators
}  // namespace hcorepp
