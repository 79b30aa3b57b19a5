// segmenter.hpp -- C++ host API with the reference Segmenter's method names.
//
// lsseg::Segmenter stands where the reference's Segmenter class stands
// (code/gpuseg/code/Segmenter.hpp:40-73): the application loads a volume, places
// a seed, asks for updates (numSubSteps iterations per request), polls the
// active count and the stop condition, and takes the binary segmentation.  It
// is a thin layer over the C ABI in lsseg.h; all device work happens there.
//
// Differences from the reference class, by design:
//   - parameters are members (Parameters) instead of global XML-backed
//     content::Parameter<float> handles (Segmenter.cpp:34-40);
//   - failures throw lsseg::Error instead of asserting;
//   - there are no OpenGL textures (the renderer is not part of this engine).
#ifndef LSSEG_SEGMENTER_HPP
#define LSSEG_SEGMENTER_HPP

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "lsseg/lsseg.h"

namespace lsseg {

class Error : public std::runtime_error {
public:
    Error(lsseg_status status, const std::string& what) : std::runtime_error(what), status_(status) {}
    lsseg_status status() const { return status_; }

private:
    lsseg_status status_;
};

// VolumeDesc (VolumeDesc.hpp:12-38) restricted to what the segmentation path
// reads; isFloat selects 32-bit float sources (an extension).
struct VolumeDesc {
    void* volumeData = nullptr;
    unsigned int numVoxelsX = 0, numVoxelsY = 0, numVoxelsZ = 0;
    float zAnisotropy = 1.0f;
    float minValue = 0.0f, maxValue = 0.0f;
    unsigned char numBytesPerVoxel = 1;
    bool isSigned = false;
    bool isFloat = false;
    // phi stored as the reference's fixed-point signed char (LEVEL_SET_CHAR,
    // Config.hpp:72-80) instead of FP32 (an extension of VolumeDesc)
    bool fixedPointPhi = false;
};

// The Segmenter system of a parameter file (bin/defaults/UnsignedChar.par).
struct Parameters {
    float growLevel = 150.0f;           // T
    float growWindow = 50.0f;           // eps
    float timeStep = 0.2f;              // dt
    float curvatureInfluence = 0.3f;    // alpha
    int numSubSteps = 10;
    float convergenceThreshold = 0.0f;
    int mode = LSSEG_MODE_REF;
};

class Segmenter {
public:
    Segmenter();
    ~Segmenter();
    Segmenter(const Segmenter&) = delete;
    Segmenter& operator=(const Segmenter&) = delete;

    // Segmenter.cpp:110-285; returns the padded description (no host data)
    VolumeDesc LoadVolume(const VolumeDesc& volumeDesc);
    void UnloadVolume();

    // Segmenter.cpp:1216-1337: new[]'d unpadded int8 {0,1}; the caller delete[]s
    // volumeData (as Engine.cpp:317 does)
    VolumeDesc GetSaveableSegmentation();

    // Segmenter.cpp:491-595
    void InitializeSeed(const float seedCoordinates[3], unsigned int sphereSize);
    // benchmark-style initialisation from a signed distance phi0 (no reference
    // counterpart): active set = {|phi0| < band}
    void InitializeLevelSet(const float* phi0, float band);
    void ClearCurrentSegmentation();    // :597-640
    void ClearAllSegmentations();       // :641-657
    void FreezeCurrentSegmentation();   // :659-679
    // :357-478: growLevel := mean, growWindow := sample standard deviation of
    // the source inside the current segmentation (export byte <= 0)
    void CalculateSegmentationParameters();

    // Segmenter.cpp:759-1051: runs numSubSteps iterations if the segmentation
    // is in progress; a change of growLevel, growWindow or curvatureInfluence
    // since the last request restarts the active set from the spatial gradient
    // (:847-876, 1371-1374)
    void RequestUpdateSegmentation(double timeDeltaSeconds);
    void RequestPlaySegmentation();
    void RequestPauseSegmentation();

    bool IsSegmentationInitialized() const;
    bool IsSegmentationFinished() const;   // :1053-1056
    bool IsSegmentationInProgress() const;

    size_t GetActiveElementCount() const;
    unsigned int GetNumIterations() const { return numIterations_; }
    VolumeDesc GetVolumeDesc() const { return paddedDesc_; }

    Parameters& GetParameters() { return params_; }
    const Parameters& GetParameters() const { return params_; }
    lsseg_ctx* GetContext() { return ctx_; }
    const lsseg_stats& GetLastStats() const { return lastStats_; }

private:
    void check(lsseg_status st, const char* what) const;
    bool ShouldInitializeActiveElements() const;
    bool ShouldUpdateSegmentation() const;

    lsseg_ctx* ctx_ = nullptr;
    VolumeDesc unpaddedDesc_{};
    VolumeDesc paddedDesc_{};
    Parameters params_{};
    float targetPrevious_ = 0.0f, windowPrevious_ = 0.0f, curvaturePrevious_ = 0.0f;
    bool initialized_ = false;
    bool paused_ = false;
    size_t numActive_ = 0;
    unsigned int numIterations_ = 0;
    lsseg_stats lastStats_{};
};

}  // namespace lsseg

#endif
