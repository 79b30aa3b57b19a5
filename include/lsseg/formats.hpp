// formats.hpp -- the reference's volume and parameter files, for callers of
// lsseg::Segmenter that open the reference's projects unchanged.
//
//   - raw volumes: VolumeLoader::LoadVolumeRaw (code/gpuseg/code/src/VolumeLoader.cpp:275-400);
//   - parameter files (*.par, the ParameterManager XML of bin/parameters and
//     bin/defaults): the "Segmenter" system's growLevel, growWindow, timeStep,
//     curvatureInfluence, numSubSteps, convergenceThreshold
//     (Segmenter.cpp:34-40);
//   - the per-type default parameter file chosen by the volume's maximum
//     (Engine.cpp:196-216);
//   - projects (*.gproj): MainWindow::openProject (src/widgets/MainWindow.cpp:692-725).
//
// DICOM series (VolumeLoader.cpp, the other loader) are out of scope.
#ifndef LSSEG_FORMATS_HPP
#define LSSEG_FORMATS_HPP

#include <string>

#include "lsseg/segmenter.hpp"

namespace lsseg {

// VolumeFileDesc (code/gpuseg/code/VolumeFileDesc.hpp)
struct VolumeFileDesc {
    std::string fileName;
    unsigned int numVoxelsX = 0, numVoxelsY = 0, numVoxelsZ = 0;
    unsigned char numBytesPerVoxel = 1;   // 1, 2 or 4; 0 marks a DICOM directory
    bool isSigned = false;
    float zAnisotropy = 1.0f;
    int upDirection[3] = {0, 0, 1};
};

// Reads nx * ny * nz voxels of numBytesPerVoxel bytes (x fastest) and their
// minimum and maximum as floats.  volumeData is new[]'d (unsigned char[]): the
// caller delete[]s it.  Throws lsseg::Error on a short or missing file.
VolumeDesc LoadVolumeRaw(const VolumeFileDesc& file);

// The Segmenter system of a parameter file; parameters it does not name keep
// their values from `base`.  Throws lsseg::Error if the file cannot be read.
Parameters LoadParameters(const std::string& parFile, const Parameters& base = Parameters());

// The default parameter set Engine::LoadVolume picks for a volume whose maximum
// is maxValue (bin/defaults/UnsignedChar.par, UnsignedShort.par,
// UnsignedShortHDR.par); throws above 65535.
Parameters DefaultParameters(float maxValue);

struct Project {
    VolumeFileDesc volume;        // fileName resolved against the root
    std::string parametersFile;   // resolved against the root; empty: defaults by type
};

// Reads a .gproj.  Its dataFile and parametersFile are relative to the
// application's working directory (the reference runs from bin/); `root`
// names that directory, and defaults to the parent of the project file's
// directory (bin/ for bin/projects/X.gproj).
Project LoadProject(const std::string& gprojFile, const std::string& root = std::string());

}  // namespace lsseg

#endif
