// Drop-in for hcorepp::operators::CompressionParameters
// (reference include/hcorepp/operators/helpers/CompressionParameters.hpp:19-128).
#pragma once

#include <hcorepp/common/Definitions.hpp>
#include <hcorepp_amd.h>

namespace hcorepp {
namespace operators {

class CompressionParameters {
public:
    CompressionParameters(double aAccuracy = 1e-4, bool aUseTrmm = false, bool aUseUngqr = true,
                          bool aTruncatedSvd = false, size_t aFixedRank = 0,
                          common::CompressionType aOpType = common::CompressionType::LAPACK_GESDD)
        : mUseTrmm(aUseTrmm), mUseUngqr(aUseUngqr), mTruncatedSvd(aTruncatedSvd), mFixedRank(aFixedRank),
          mOpType(aOpType), mAccuracy(aAccuracy) {}
    bool GetTrmm() const { return mUseTrmm; }
    bool GetUngqr() const { return mUseUngqr; }
    bool GetTruncatedSvd() const { return mTruncatedSvd; }
    size_t GetFixedRank() const { return mFixedRank; }
    common::CompressionType GetOperationType() const { return mOpType; }
    double GetAccuracy() const { return mAccuracy; }
    hcorepp_cparams ToC() const {
        return hcorepp_cparams{mAccuracy, (int32_t) mUseTrmm, (int32_t) mUseUngqr, (int32_t) mTruncatedSvd,
                               (int32_t) mOpType, (int64_t) mFixedRank};
    }

private:
    bool mUseTrmm, mUseUngqr, mTruncatedSvd;
    size_t mFixedRank;
    common::CompressionType mOpType;
    double mAccuracy;
};

}  // namespace operators
}  // namespace hcorepp
