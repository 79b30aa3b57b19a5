// segmenter.cpp -- lsseg::Segmenter over the C ABI (see include/lsseg/segmenter.hpp).
#include "lsseg/segmenter.hpp"

#include <cstring>
#include <string>

namespace lsseg {

namespace {
// MINIMUM_SEGMENTATION_ITERATIONS (Config.hpp:45)
constexpr unsigned int kMinimumIterations = 500;

int32_t dtype_of(const VolumeDesc& d)
{
    if (d.isFloat && d.numBytesPerVoxel == 4) return LSSEG_F32;
    switch (d.numBytesPerVoxel) {
    case 1: return d.isSigned ? LSSEG_I8 : LSSEG_U8;
    case 2: return d.isSigned ? LSSEG_I16 : LSSEG_U16;
    case 4: return d.isSigned ? LSSEG_I32 : LSSEG_U32;
    default: return -1;
    }
}
}  // namespace

Segmenter::Segmenter() = default;

Segmenter::~Segmenter()
{
    if (ctx_) lsseg_destroy(ctx_);
}

void Segmenter::check(lsseg_status st, const char* what) const
{
    if (st != LSSEG_OK) throw Error(st, std::string(what) + ": " + lsseg_last_error(ctx_));
}

VolumeDesc Segmenter::LoadVolume(const VolumeDesc& volumeDesc)
{
    UnloadVolume();
    const int32_t dt = dtype_of(volumeDesc);
    if (dt < 0) throw Error(LSSEG_ERR_UNSUPPORTED, "LoadVolume: unsupported voxel type");
    lsseg_volume_desc d{volumeDesc.numVoxelsX, volumeDesc.numVoxelsY, volumeDesc.numVoxelsZ, dt,
                        volumeDesc.zAnisotropy, volumeDesc.fixedPointPhi ? LSSEG_PHI_I8 : LSSEG_PHI_F32};
    lsseg_ctx* c = nullptr;
    const lsseg_status st = lsseg_create(&d, volumeDesc.volumeData, &c);
    if (st != LSSEG_OK) throw Error(st, std::string("LoadVolume: ") + lsseg_last_error(nullptr));
    ctx_ = c;
    unpaddedDesc_ = volumeDesc;
    unpaddedDesc_.volumeData = nullptr;
    uint32_t X = 0, Y = 0, Z = 0;
    check(lsseg_get_dims(ctx_, &X, &Y, &Z), "LoadVolume");
    paddedDesc_ = unpaddedDesc_;
    paddedDesc_.numVoxelsX = X;
    paddedDesc_.numVoxelsY = Y;
    paddedDesc_.numVoxelsZ = Z;
    initialized_ = false;
    numActive_ = 0;
    numIterations_ = 0;
    return paddedDesc_;
}

void Segmenter::UnloadVolume()
{
    if (ctx_) lsseg_destroy(ctx_);
    ctx_ = nullptr;
    initialized_ = false;
    numActive_ = 0;
    numIterations_ = 0;
}

VolumeDesc Segmenter::GetSaveableSegmentation()
{
    if (!ctx_) throw Error(LSSEG_ERR_STATE, "GetSaveableSegmentation: no volume loaded");
    const size_t n = (size_t)unpaddedDesc_.numVoxelsX * unpaddedDesc_.numVoxelsY * unpaddedDesc_.numVoxelsZ;
    signed char* buf = new signed char[n];
    const lsseg_status st = lsseg_get_mask(ctx_, reinterpret_cast<uint8_t*>(buf));
    if (st != LSSEG_OK) {
        delete[] buf;
        check(st, "GetSaveableSegmentation");
    }
    VolumeDesc out = unpaddedDesc_;
    out.volumeData = buf;
    out.numBytesPerVoxel = 1;
    out.isSigned = true;
    out.isFloat = false;
    out.minValue = 0.0f;
    out.maxValue = 0.0f;
    for (size_t i = 0; i < n; ++i)
        if (buf[i]) {
            out.maxValue = 1.0f;
            break;
        }
    return out;
}

void Segmenter::InitializeSeed(const float seedCoordinates[3], unsigned int sphereSize)
{
    if (!ctx_) throw Error(LSSEG_ERR_STATE, "InitializeSeed: no volume loaded");
    check(lsseg_seed_sphere(ctx_, seedCoordinates[0], seedCoordinates[1], seedCoordinates[2], sphereSize),
          "InitializeSeed");
    uint64_t n = 0;
    check(lsseg_get_active_count(ctx_, &n), "InitializeSeed");
    numActive_ = (size_t)n;
    initialized_ = true;
    numIterations_ = 0;
    targetPrevious_ = params_.growLevel;
    windowPrevious_ = params_.growWindow;
    curvaturePrevious_ = params_.curvatureInfluence;
}

void Segmenter::InitializeLevelSet(const float* phi0, float band)
{
    if (!ctx_) throw Error(LSSEG_ERR_STATE, "InitializeLevelSet: no volume loaded");
    check(lsseg_set_phi(ctx_, phi0, band), "InitializeLevelSet");
    uint64_t n = 0;
    check(lsseg_get_active_count(ctx_, &n), "InitializeLevelSet");
    numActive_ = (size_t)n;
    initialized_ = true;
    numIterations_ = 0;
    targetPrevious_ = params_.growLevel;
    windowPrevious_ = params_.growWindow;
    curvaturePrevious_ = params_.curvatureInfluence;
}

void Segmenter::ClearCurrentSegmentation()
{
    if (!ctx_) return;
    check(lsseg_clear(ctx_, 0), "ClearCurrentSegmentation");
    initialized_ = false;
    numActive_ = 0;
    numIterations_ = 0;
}

void Segmenter::ClearAllSegmentations()
{
    if (!ctx_) return;
    check(lsseg_clear(ctx_, 1), "ClearAllSegmentations");
    initialized_ = false;
    numActive_ = 0;
    numIterations_ = 0;
}

void Segmenter::FreezeCurrentSegmentation()
{
    if (!ctx_) throw Error(LSSEG_ERR_STATE, "FreezeCurrentSegmentation: no volume loaded");
    check(lsseg_freeze(ctx_), "FreezeCurrentSegmentation");
}

void Segmenter::CalculateSegmentationParameters()
{
    if (!ctx_) throw Error(LSSEG_ERR_STATE, "CalculateSegmentationParameters: no volume loaded");
    float T = 0.0f, eps = 0.0f;
    check(lsseg_auto_parameters(ctx_, &T, &eps, nullptr), "CalculateSegmentationParameters");
    params_.growLevel = T;
    params_.growWindow = eps;
}

bool Segmenter::ShouldInitializeActiveElements() const
{
    // Segmenter.cpp:1371-1374
    return targetPrevious_ != params_.growLevel || windowPrevious_ != params_.growWindow ||
           curvaturePrevious_ != params_.curvatureInfluence;
}

bool Segmenter::ShouldUpdateSegmentation() const
{
    // Segmenter.cpp:1358-1361
    return ((IsSegmentationInitialized() && !IsSegmentationFinished()) ||
            (initialized_ && ShouldInitializeActiveElements())) &&
           !paused_;
}

void Segmenter::RequestUpdateSegmentation(double)
{
    if (!ctx_ || !ShouldUpdateSegmentation()) return;
    lsseg_params p{params_.growLevel, params_.growWindow, params_.curvatureInfluence, params_.timeStep,
                   params_.mode};
    int steps = params_.numSubSteps;
    memset(&lastStats_, 0, sizeof(lastStats_));
    if (steps > 0 && ShouldInitializeActiveElements()) {
        // first sub-step: update, then restart the active set from the spatial
        // gradient of the new phi (Segmenter.cpp:847-876, 1520-1526)
        check(lsseg_iterate_reinit(ctx_, &p), "RequestUpdateSegmentation");
        targetPrevious_ = params_.growLevel;
        windowPrevious_ = params_.growWindow;
        curvaturePrevious_ = params_.curvatureInfluence;
        numIterations_ = 1;
        --steps;
    }
    if (steps > 0) {
        check(lsseg_iterate(ctx_, &p, steps, &lastStats_), "RequestUpdateSegmentation");
        numIterations_ += (unsigned int)steps;
    }
    uint64_t n = 0;
    check(lsseg_get_active_count(ctx_, &n), "RequestUpdateSegmentation");
    numActive_ = (size_t)n;
}

void Segmenter::RequestPlaySegmentation() { paused_ = false; }
void Segmenter::RequestPauseSegmentation() { paused_ = true; }

bool Segmenter::IsSegmentationInitialized() const { return initialized_; }

bool Segmenter::IsSegmentationFinished() const
{
    // Segmenter.cpp:1053-1056
    return (double)numActive_ <= (double)params_.convergenceThreshold && numIterations_ > kMinimumIterations;
}

bool Segmenter::IsSegmentationInProgress() const { return ShouldUpdateSegmentation(); }

size_t Segmenter::GetActiveElementCount() const { return numActive_; }

}  // namespace lsseg
