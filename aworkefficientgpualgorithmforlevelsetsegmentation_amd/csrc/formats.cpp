// formats.cpp -- raw volumes, parameter files and projects of the reference
// application (see include/lsseg/formats.hpp for the reference lines each
// function follows).  Host code only.
#include "lsseg/formats.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <regex>
#include <sstream>

namespace lsseg {

namespace {

std::string read_text(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(LSSEG_ERR_INVALID_ARG, "cannot open " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

// attribute name -> value of one start tag's text
std::map<std::string, std::string> attributes(const std::string& tag)
{
    std::map<std::string, std::string> out;
    static const std::regex attr(R"re(([A-Za-z_][A-Za-z0-9_]*)\s*=\s*"([^"]*)")re");
    for (std::sregex_iterator it(tag.begin(), tag.end(), attr), end; it != end; ++it) out[(*it)[1]] = (*it)[2];
    return out;
}

// the start tag beginning at `pos` ('<' ... '>')
std::string start_tag(const std::string& text, size_t pos)
{
    const size_t close = text.find('>', pos);
    if (close == std::string::npos) throw Error(LSSEG_ERR_INVALID_ARG, "unterminated tag");
    return text.substr(pos, close - pos + 1);
}

std::string dirname(const std::string& path)
{
    const size_t s = path.find_last_of('/');
    return s == std::string::npos ? std::string(".") : (s == 0 ? std::string("/") : path.substr(0, s));
}

std::string join(const std::string& root, const std::string& rel)
{
    if (rel.empty() || rel[0] == '/') return rel;
    std::string r = rel;
    for (char& ch : r)
        if (ch == '\\') ch = '/';
    return root + "/" + r;
}

template <typename T>
void minmax(const unsigned char* data, size_t count, float* lo, float* hi)
{
    // VolumeLoader.cpp:305-384: starts from 65535 / -1
    float mn = 65535.0f, mx = -1.0f;
    const T* v = reinterpret_cast<const T*>(data);
    for (size_t i = 0; i < count; ++i) {
        const float x = static_cast<float>(v[i]);
        mn = x < mn ? x : mn;
        mx = x > mx ? x : mx;
    }
    *lo = mn;
    *hi = mx;
}

}  // namespace

VolumeDesc LoadVolumeRaw(const VolumeFileDesc& file)
{
    const size_t count = (size_t)file.numVoxelsX * file.numVoxelsY * file.numVoxelsZ;
    const int b = file.numBytesPerVoxel;
    if (count == 0 || (b != 1 && b != 2 && b != 4))
        throw Error(LSSEG_ERR_INVALID_ARG, "raw volumes need positive dims and 1, 2 or 4 bytes per voxel");
    const size_t bytes = count * (size_t)b;
    unsigned char* data = new unsigned char[bytes];
    FILE* f = std::fopen(file.fileName.c_str(), "rb");
    const size_t got = f ? std::fread(data, 1, bytes, f) : 0;
    if (f) std::fclose(f);
    if (got != bytes) {
        delete[] data;
        throw Error(LSSEG_ERR_INVALID_ARG, "cannot read " + std::to_string(bytes) + " bytes from " + file.fileName);
    }
    VolumeDesc d;
    d.volumeData = data;
    d.numVoxelsX = file.numVoxelsX;
    d.numVoxelsY = file.numVoxelsY;
    d.numVoxelsZ = file.numVoxelsZ;
    d.zAnisotropy = file.zAnisotropy;
    d.numBytesPerVoxel = (unsigned char)b;
    d.isSigned = file.isSigned;
    d.isFloat = false;
    if (b == 1) file.isSigned ? minmax<int8_t>(data, count, &d.minValue, &d.maxValue)
                              : minmax<uint8_t>(data, count, &d.minValue, &d.maxValue);
    else if (b == 2) file.isSigned ? minmax<int16_t>(data, count, &d.minValue, &d.maxValue)
                                   : minmax<uint16_t>(data, count, &d.minValue, &d.maxValue);
    else file.isSigned ? minmax<int32_t>(data, count, &d.minValue, &d.maxValue)
                       : minmax<uint32_t>(data, count, &d.minValue, &d.maxValue);
    return d;
}

Parameters LoadParameters(const std::string& parFile, const Parameters& base)
{
    const std::string text = read_text(parFile);
    Parameters p = base;
    // the <System name="Segmenter"> element
    size_t pos = 0, sys_begin = std::string::npos;
    while ((pos = text.find("<System", pos)) != std::string::npos) {
        auto at = attributes(start_tag(text, pos));
        if (at["name"] == "Segmenter") {
            sys_begin = pos;
            break;
        }
        ++pos;
    }
    if (sys_begin == std::string::npos) throw Error(LSSEG_ERR_INVALID_ARG, parFile + ": no Segmenter system");
    const size_t sys_end = text.find("</System>", sys_begin);
    for (pos = text.find("<Parameter", sys_begin); pos != std::string::npos && pos < sys_end;
         pos = text.find("<Parameter", pos + 1)) {
        if (text.compare(pos, 11, "<Parameters") == 0) continue;
        auto at = attributes(start_tag(text, pos));
        const size_t vpos = text.find("<ValueFloat", pos);
        const size_t pend = text.find("</Parameter>", pos);
        if (vpos == std::string::npos || vpos > pend) continue;
        const float v = std::strtof(attributes(start_tag(text, vpos))["value"].c_str(), nullptr);
        const std::string& name = at["name"];
        if (name == "growLevel") p.growLevel = v;
        else if (name == "growWindow") p.growWindow = v;
        else if (name == "timeStep") p.timeStep = v;
        else if (name == "curvatureInfluence") p.curvatureInfluence = v;
        else if (name == "numSubSteps") p.numSubSteps = (int)v;
        else if (name == "convergenceThreshold") p.convergenceThreshold = v;
    }
    return p;
}

Parameters DefaultParameters(float maxValue)
{
    // Engine.cpp:196-216 with the values of bin/defaults/*.par
    Parameters p;
    p.timeStep = 0.2f;
    p.curvatureInfluence = 0.3f;
    p.numSubSteps = 10;
    p.convergenceThreshold = 0.0f;
    if (maxValue <= 255.0f) {
        p.growLevel = 150.0f;
        p.growWindow = 50.0f;
    } else if (maxValue <= 4095.0f) {
        p.growLevel = 790.0f;
        p.growWindow = 350.0f;
    } else if (maxValue <= 65535.0f) {
        p.growLevel = 15000.0f;
        p.growWindow = 5000.0f;
    } else {
        throw Error(LSSEG_ERR_UNSUPPORTED, "no default parameters for a maximum above 65535");
    }
    return p;
}

Project LoadProject(const std::string& gprojFile, const std::string& rootIn)
{
    const std::string text = read_text(gprojFile);
    const size_t pos = text.find("<GPUSegProject");
    if (pos == std::string::npos) throw Error(LSSEG_ERR_INVALID_ARG, gprojFile + ": no GPUSegProject element");
    auto at = attributes(start_tag(text, pos));
    const std::string root = rootIn.empty() ? dirname(dirname(gprojFile)) : rootIn;
    Project pr;
    pr.volume.fileName = join(root, at["dataFile"]);
    pr.volume.isSigned = std::atoi(at["isSigned"].c_str()) != 0;
    pr.volume.numBytesPerVoxel = (unsigned char)std::atoi(at["numBytesPerVoxel"].c_str());
    pr.volume.numVoxelsX = (unsigned)std::atoi(at["numVoxelsX"].c_str());
    pr.volume.numVoxelsY = (unsigned)std::atoi(at["numVoxelsY"].c_str());
    pr.volume.numVoxelsZ = (unsigned)std::atoi(at["numVoxelsZ"].c_str());
    pr.volume.upDirection[0] = std::atoi(at["upDirectionX"].c_str());
    pr.volume.upDirection[1] = std::atoi(at["upDirectionY"].c_str());
    pr.volume.upDirection[2] = std::atoi(at["upDirectionZ"].c_str());
    pr.volume.zAnisotropy = at.count("zAnisotropy") ? std::strtof(at["zAnisotropy"].c_str(), nullptr) : 1.0f;
    pr.parametersFile = at["parametersFile"].empty() ? std::string() : join(root, at["parametersFile"]);
    return pr;
}

}  // namespace lsseg
