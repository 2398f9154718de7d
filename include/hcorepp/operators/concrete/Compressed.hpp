// Drop-in for hcorepp::operators::CompressedTile<T>
// (reference include/hcorepp/operators/concrete/Compressed.hpp:29-258,
// src/operators/concrete/Compressed.cpp:19-73): one device buffer
// [U rows x maxRank (ld rows) | V maxRank x cols (ld rank)], maxRank = rank at construction
// as in the reference, so the rank of a C tile never grows past its initial rank.
#pragma once

#include <algorithm>

#include <hcorepp/operators/interface/Tile.hpp>

namespace hcorepp {
namespace operators {

template <typename T>
class CompressedTile : public Tile<T> {
public:
    // U (rows x rank) and V (rank x cols) given separately (Compressed.cpp:19-47).
    CompressedTile(size_t aNumOfRows, size_t aNumOfCols, T *apDataU, T *apDataV, size_t aLeadingDim, size_t aRank,
                   blas::Layout aLayout, const kernels::RunContext &aContext) {
        (void) aLayout;
        Init(aNumOfRows, aNumOfCols, aLeadingDim, aRank, aRank);
        CopyIn(apDataU, apDataV, aContext);
    }
    // apData holds U (rows x rank) followed by V (rank x cols) (Compressed.cpp:49-73).
    CompressedTile(size_t aNumOfRows, size_t aNumOfCols, T *apData, size_t aLeadingDim, size_t aRank,
                   blas::Layout aLayout, const kernels::RunContext &aContext) {
        (void) aLayout;
        Init(aNumOfRows, aNumOfCols, aLeadingDim, aRank, aRank);
        if (apData) CopyIn(apData, apData + aNumOfRows * aRank, aContext);
    }
    CompressedTile(size_t aNumOfRows, size_t aNumOfCols, T *apData, size_t aLeadingDim, size_t aRank,
                   const kernels::RunContext &aContext)
        : CompressedTile(aNumOfRows, aNumOfCols, apData, aLeadingDim, aRank, blas::Layout::ColMajor, aContext) {}
    // Extension: capacity aMaxRank >= aRank, so recompression may raise the rank up to it.
    CompressedTile(size_t aNumOfRows, size_t aNumOfCols, T *apData, size_t aLeadingDim, size_t aRank,
                   size_t aMaxRank, const kernels::RunContext &aContext) {
        Init(aNumOfRows, aNumOfCols, aLeadingDim, aRank, std::max(aRank, aMaxRank));
        if (apData) CopyIn(apData, apData + aNumOfRows * aRank, aContext);
    }
    // Wraps caller-owned device memory described by a TileMetadata (PackTile, Compressed.cpp:781-799).
    CompressedTile(const TileMetadata &aMeta, T *apDeviceData) {
        this->mDesc = hcorepp_tile{(int64_t) aMeta.mNumOfRows, (int64_t) aMeta.mNumOfCols,
                                   (int64_t) aMeta.mMatrixRank, (int64_t) aMeta.mMaxRank,
                                   (int64_t) aMeta.mLeadingDimension, 0, HCOREPP_COMPRESSED, apDeviceData};
    }
    T *GetUMatrix() const { return this->GetData(); }
    T *GetVMatrix() const { return this->GetData() + this->mDesc.rows * this->mDesc.max_rank; }
    size_t GetTileRank() const { return (size_t) this->mDesc.rank; }
    size_t GetMaxRank() const { return (size_t) this->mDesc.max_rank; }
    size_t GetULeadingDim() const { return (size_t) this->mDesc.rows; }
    size_t GetVLeadingDim() const { return (size_t) this->mDesc.rank; }
    T *GetTileSubMatrix(size_t aIndex) const {
        if (aIndex > 1) throw std::invalid_argument("CompressedTile::GetTileSubMatrix: index must be 0 or 1");
        return aIndex == 0 ? GetUMatrix() : GetVMatrix();
    }
    TileType GetTileType() const { return COMPRESSED; }
    TileMetadata UnPackTile() const {
        return TileMetadata(this->GetNumOfRows(), this->GetNumOfCols(), GetTileRank(), GetMaxRank(),
                            this->GetLeadingDim(), blas::Layout::ColMajor, COMPRESSED);
    }

private:
    void Init(size_t aRows, size_t aCols, size_t aLd, size_t aRank, size_t aMaxRank) {
        this->mDesc = hcorepp_tile{(int64_t) aRows, (int64_t) aCols, (int64_t) aRank, (int64_t) aMaxRank,
                                   (int64_t) aLd, 0, HCOREPP_COMPRESSED, nullptr};
        this->Allocate((aRows + aCols) * std::max<size_t>(aMaxRank, 1));
    }
    void CopyIn(const T *apU, const T *apV, const kernels::RunContext &aContext) {
        const size_t m = (size_t) this->mDesc.rows, n = (size_t) this->mDesc.cols, r = (size_t) this->mDesc.rank;
        if (apU) kernels::ThrowIfError(hcorepp_memcpy(aContext.Handle(), GetUMatrix(), apU, m * r * sizeof(T)));
        if (apV) kernels::ThrowIfError(hcorepp_memcpy(aContext.Handle(), GetVMatrix(), apV, r * n * sizeof(T)));
    }
};

}  // namespace operators
}  // namespace hcorepp
