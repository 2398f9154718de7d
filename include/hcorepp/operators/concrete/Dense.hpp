// Drop-in for hcorepp::operators::DenseTile<T>
// (reference include/hcorepp/operators/concrete/Dense.hpp:56-95, src/operators/concrete/Dense.cpp:19-29).
#pragma once

#include <algorithm>

#include <hcorepp/operators/interface/Tile.hpp>

namespace hcorepp {
namespace operators {

template <typename T>
class DenseTile : public Tile<T> {
public:
    // Copies ld x cols elements from aPdata (host or device memory; may be null for a zero tile).
    // With aMemoryOwnership = false the tile wraps device memory aPdata without copying.
    DenseTile(size_t aNumOfRows, size_t aNumOfCols, T *aPdata, size_t aLeadingDim, blas::Layout aLayout,
              const kernels::RunContext &aContext, bool aMemoryOwnership = true) {
        (void) aLayout;
        (void) aContext;
        // a dense tile carries no rank (its metadata reports 0, as the reference's does)
        this->mDesc = hcorepp_tile{(int64_t) aNumOfRows, (int64_t) aNumOfCols, 0, 0, (int64_t) aLeadingDim, 0,
                                   HCOREPP_DENSE, nullptr};
        if (!aMemoryOwnership) {
            this->mDesc.data = aPdata;
            return;
        }
        this->Allocate(aLeadingDim * aNumOfCols);
        if (aPdata)
            kernels::ThrowIfError(
                    hcorepp_memcpy(aContext.Handle(), this->mDesc.data, aPdata, aLeadingDim * aNumOfCols * sizeof(T)));
    }
    DenseTile(size_t aNumOfRows, size_t aNumOfCols, T *aPdata, size_t aLeadingDim,
              const kernels::RunContext &aContext)
        : DenseTile(aNumOfRows, aNumOfCols, aPdata, aLeadingDim, blas::Layout::ColMajor, aContext) {}
    // Non-owning tile over caller memory described by aMeta (PackTile, Dense.cpp:136-142).
    DenseTile(const TileMetadata &aMeta, T *apDeviceData) {
        this->mDesc = hcorepp_tile{(int64_t) aMeta.mNumOfRows, (int64_t) aMeta.mNumOfCols, 0, 0,
                                   (int64_t) aMeta.mLeadingDimension, 0, HCOREPP_DENSE, apDeviceData};
    }
    T *GetTileSubMatrix(size_t aIndex) const {
        if (aIndex != 0) throw std::invalid_argument("DenseTile::GetTileSubMatrix: index must be 0");
        return this->GetData();
    }
    TileType GetTileType() const { return DENSE; }
};

}  // namespace operators
}  // namespace hcorepp
