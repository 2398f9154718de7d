// Drop-in for the pool MemoryUnit / MemoryHandler singleton
// (reference include/hcorepp/data-units/memory-handlers/pool/MemoryHandler.hpp:22-157).
// Workspace lives in the RunContext arena here, so a MemoryUnit carries no memory of its
// own; it exists so that calls written against the reference signature compile unchanged.
#pragma once

#include <cstddef>
#include <vector>

namespace hcorepp {
namespace dataunits {

template <typename T>
class MemoryUnit {
public:
    bool IsInitialized() const { return true; }
    void Initialize(size_t) {}
    void Reset() {}
    void FreeAllocations() {}
};

template <typename T>
class MemoryHandler {
public:
    static MemoryHandler &GetInstance() {
        static MemoryHandler h;
        return h;
    }
    MemoryUnit<T> &GetMemoryUnit(size_t aIdx = 0) {
        if (aIdx >= mUnits.size()) mUnits.resize(aIdx + 1);
        return mUnits[aIdx];
    }
    void FreeAllocations() {}
    static void DestroyInstance() {}

private:
    std::vector<MemoryUnit<T>> mUnits = std::vector<MemoryUnit<T>>(1);
};

}  // namespace dataunits
}  // namespace hcorepp
