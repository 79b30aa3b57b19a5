// segmenter_demo.cpp -- drives lsseg::Segmenter the way the reference
// application drives its Segmenter (Engine.cpp:179-257, 366-424, 732-742,
// 309-318): load a raw volume, place a seed, request updates until the
// segmentation finishes or a request budget runs out, save the mask.
//
// usage: segmenter_demo <raw> <nx> <ny> <nz> <bytes> <signed> <cx> <cy> <cz> <radius>
//                       <T> <eps> <dt> <alpha> <substeps> <max_requests> <mask_out>
//        segmenter_demo --parse <project.gproj>
//            prints the project and its parameters (no device work)
//        segmenter_demo --project <project.gproj> <cx> <cy> <cz> <radius> <max_requests> <mask_out> <auto>
//            opens a reference project (MainWindow.cpp:692-725, Engine.cpp:179-257);
//            auto = 1 runs CalculateSegmentationParameters after the seed
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "lsseg/formats.hpp"
#include "lsseg/segmenter.hpp"

namespace {

void print_parameters(const lsseg::Parameters& p)
{
    std::printf("growLevel %.9g\ngrowWindow %.9g\ntimeStep %.9g\ncurvatureInfluence %.9g\nnumSubSteps %d\n"
                "convergenceThreshold %.9g\n",
                p.growLevel, p.growWindow, p.timeStep, p.curvatureInfluence, p.numSubSteps, p.convergenceThreshold);
}

lsseg::Parameters project_parameters(const lsseg::Project& pr, const lsseg::VolumeDesc& vol)
{
    // Engine.cpp:196-226: the project's file, else the default for the type
    return pr.parametersFile.empty() ? lsseg::DefaultParameters(vol.maxValue)
                                     : lsseg::LoadParameters(pr.parametersFile);
}

int run_project(int argc, char** argv)
{
    const bool parse_only = std::strcmp(argv[1], "--parse") == 0;
    if ((parse_only && argc != 3) || (!parse_only && argc != 10)) {
        std::fprintf(stderr, "usage: %s --parse gproj | --project gproj cx cy cz r max_requests mask_out auto\n",
                     argv[0]);
        return 2;
    }
    try {
        const lsseg::Project pr = lsseg::LoadProject(argv[2]);
        lsseg::VolumeDesc vol = lsseg::LoadVolumeRaw(pr.volume);
        lsseg::Parameters params = project_parameters(pr, vol);
        if (parse_only) {
            std::printf("dataFile %s\nparametersFile %s\nnumVoxels %u %u %u\nnumBytesPerVoxel %d\nisSigned %d\n"
                        "zAnisotropy %.9g\nupDirection %d %d %d\nminValue %.9g\nmaxValue %.9g\n",
                        pr.volume.fileName.c_str(), pr.parametersFile.c_str(), pr.volume.numVoxelsX,
                        pr.volume.numVoxelsY, pr.volume.numVoxelsZ, pr.volume.numBytesPerVoxel,
                        pr.volume.isSigned ? 1 : 0, pr.volume.zAnisotropy, pr.volume.upDirection[0],
                        pr.volume.upDirection[1], pr.volume.upDirection[2], vol.minValue, vol.maxValue);
            print_parameters(params);
            delete[] static_cast<unsigned char*>(vol.volumeData);
            return 0;
        }
        const size_t count = (size_t)vol.numVoxelsX * vol.numVoxelsY * vol.numVoxelsZ;
        lsseg::Segmenter seg;
        seg.LoadVolume(vol);
        delete[] static_cast<unsigned char*>(vol.volumeData);
        seg.GetParameters() = params;
        const float seed[3] = {(float)std::atof(argv[3]), (float)std::atof(argv[4]), (float)std::atof(argv[5])};
        seg.InitializeSeed(seed, (unsigned)std::atoi(argv[6]));
        if (std::atoi(argv[9])) seg.CalculateSegmentationParameters();
        print_parameters(seg.GetParameters());
        const int max_requests = std::atoi(argv[7]);
        int requests = 0;
        while (seg.IsSegmentationInProgress() && requests < max_requests) {
            seg.RequestUpdateSegmentation(0.0);
            ++requests;
        }
        lsseg::VolumeDesc mask = seg.GetSaveableSegmentation();
        const signed char* m = (const signed char*)mask.volumeData;
        FILE* o = std::fopen(argv[8], "wb");
        const bool ok = o && std::fwrite(m, 1, count, o) == count;
        if (o) std::fclose(o);
        delete[] m;
        if (!ok) {
            std::fprintf(stderr, "cannot write %s\n", argv[8]);
            return 2;
        }
        std::printf("requests %d iterations %u active %zu\n", requests, seg.GetNumIterations(),
                    seg.GetActiveElementCount());
    } catch (const lsseg::Error& e) {
        std::fprintf(stderr, "lsseg error %d: %s\n", (int)e.status(), e.what());
        return 1;
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc > 1 && (std::strcmp(argv[1], "--parse") == 0 || std::strcmp(argv[1], "--project") == 0))
        return run_project(argc, argv);
    if (argc != 18) {
        std::fprintf(stderr, "usage: %s raw nx ny nz bytes signed cx cy cz r T eps dt alpha substeps max_requests mask_out\n",
                     argv[0]);
        return 2;
    }
    const unsigned nx = std::atoi(argv[2]), ny = std::atoi(argv[3]), nz = std::atoi(argv[4]);
    const int bytes = std::atoi(argv[5]);
    const bool is_signed = std::atoi(argv[6]) != 0;
    const size_t count = (size_t)nx * ny * nz;
    std::vector<unsigned char> raw(count * bytes);
    FILE* f = std::fopen(argv[1], "rb");
    if (!f || std::fread(raw.data(), 1, raw.size(), f) != raw.size()) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 2;
    }
    std::fclose(f);
    try {
        lsseg::Segmenter seg;
        lsseg::VolumeDesc desc;
        desc.volumeData = raw.data();
        desc.numVoxelsX = nx;
        desc.numVoxelsY = ny;
        desc.numVoxelsZ = nz;
        desc.numBytesPerVoxel = (unsigned char)bytes;
        desc.isSigned = is_signed;
        desc.isFloat = false;
        const lsseg::VolumeDesc padded = seg.LoadVolume(desc);
        lsseg::Parameters& p = seg.GetParameters();
        p.growLevel = (float)std::atof(argv[11]);
        p.growWindow = (float)std::atof(argv[12]);
        p.timeStep = (float)std::atof(argv[13]);
        p.curvatureInfluence = (float)std::atof(argv[14]);
        p.numSubSteps = std::atoi(argv[15]);
        const float seed[3] = {(float)std::atof(argv[7]), (float)std::atof(argv[8]), (float)std::atof(argv[9])};
        seg.InitializeSeed(seed, (unsigned)std::atoi(argv[10]));
        std::printf("padded %u %u %u initial_active %zu\n", padded.numVoxelsX, padded.numVoxelsY,
                    padded.numVoxelsZ, seg.GetActiveElementCount());
        const int max_requests = std::atoi(argv[16]);
        int requests = 0;
        while (seg.IsSegmentationInProgress() && requests < max_requests) {
            seg.RequestUpdateSegmentation(0.0);
            ++requests;
        }
        lsseg::VolumeDesc mask = seg.GetSaveableSegmentation();
        size_t inside = 0;
        const signed char* m = (const signed char*)mask.volumeData;
        for (size_t i = 0; i < count; ++i) inside += m[i] ? 1 : 0;
        FILE* o = std::fopen(argv[17], "wb");
        if (!o || std::fwrite(m, 1, count, o) != count) {
            std::fprintf(stderr, "cannot write %s\n", argv[17]);
            delete[] m;
            return 2;
        }
        std::fclose(o);
        std::printf("requests %d iterations %u active %zu finished %d segmented %zu\n", requests,
                    seg.GetNumIterations(), seg.GetActiveElementCount(), seg.IsSegmentationFinished() ? 1 : 0,
                    inside);
        delete[] m;
    } catch (const lsseg::Error& e) {
        std::fprintf(stderr, "lsseg error %d: %s\n", (int)e.status(), e.what());
        return 1;
    }
    return 0;
}
