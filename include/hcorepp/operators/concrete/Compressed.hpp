// Drop-in for hcorepp::operators::CompressedTile<T>
// (reference include/hcorepp/operators/concrete/Compressed.hpp:29-258,
// src/operators/concrete/Compressed.cpp:19-146): one device buffer
// [U rows x maxRank (ld rows) | V maxRank x cols (ld rank)], maxRank = rank at construction
// as in the reference, so the rank of a C tile never grows past its initial rank.
#pragma once

#include <algorithm>

#include <hcorepp/operators/helpers/CompressionParameters.hpp>
#include <hcorepp/operators/interface/Tile.hpp>

namespace hcorepp {
namespace operators {

namespace detail {
inline hcorepp_status compress(hcorepp_ctx *c, int64_t m, int64_t n, const double *a, int64_t lda,
                               const hcorepp_cparams *p, hcorepp_tile *t) {
    return hcorepp_compress_d(c, m, n, a, lda, p, t);
}
inline hcorepp_status compress(hcorepp_ctx *c, int64_t m, int64_t n, const float *a, int64_t lda,
                               const hcorepp_cparams *p, hcorepp_tile *t) {
    return hcorepp_compress_s(c, m, n, a, lda, p, t);
}
}  // namespace detail

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
    // Compression constructor (Compressed.cpp:75-146): truncated SVD of the dense tile at apData
    // (host or device memory, ld aLeadingDim) on the device; maxRank = max(min(rows, cols) / 3, 1)
    // as in the reference; without data the rank is maxRank and the buffer is left unset.
    CompressedTile(size_t aNumOfRows, size_t aNumOfCols, T *apData, size_t aLeadingDim,
                   const CompressionParameters &aParameters, blas::Layout aLayout,
                   const kernels::RunContext &aContext) {
        (void) aLayout;
        const size_t mr = std::max<size_t>(std::min(aNumOfRows, aNumOfCols) / 3, 1);
        Init(aNumOfRows, aNumOfCols, aNumOfRows, mr, mr);
        if (!apData) return;
        const size_t bytes = aLeadingDim * aNumOfCols * sizeof(T);
        void *dev = nullptr;
        kernels::ThrowIfError(hcorepp_malloc(bytes, &dev));
        hcorepp_status st = hcorepp_memcpy(aContext.Handle(), dev, apData, bytes);
        const hcorepp_cparams p = aParameters.ToC();
        if (st == HCOREPP_OK)
            st = detail::compress(aContext.Handle(), (int64_t) aNumOfRows, (int64_t) aNumOfCols, (const T *) dev,
                                  (int64_t) aLeadingDim, &p, &this->mDesc);
        hcorepp_free(dev);
        kernels::ThrowIfError(st);
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
