// Example application of the C++ drop-in (include/isaac_mi355x.hpp): one
// persistent float volume source rendered through IsaacVisualization, the
// way example/example.cpp drives the reference.  Used by the GPU tests:
//
//   shim_example <field.raw> <n> <W> <H> <out.rgba>
//
// reads an n^3 float32 field (x fastest), renders it with the 4-point
// transfer function of the small workload (trilinear, step 0.5) and writes
// the W*H*4 RGBA8 framebuffer.
#include "isaac_json_stub.h"
#include "isaac_mi355x.hpp"

#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <tuple>
#include <vector>

struct DensitySource
{
    static const size_t featureDim = 1;
    static const size_t guardSize = 0;
    static const bool persistent = true;
    static std::string getName()
    {
        return "density";
    }
    float* ptr = nullptr;
    int64_t X = 0, Y = 0;
    void update(bool enabled, void* pointer)
    {
        (void) enabled;
        (void) pointer;
    }
    isaac::VolumeView view() const
    {
        return isaac::VolumeView{ptr, {1, X, X * Y}};
    }
};

int main(int argc, char** argv)
{
    if(argc != 6)
    {
        std::fprintf(stderr, "usage: %s field.raw n W H out.rgba\n", argv[0]);
        return 2;
    }
    const int64_t n = std::atoll(argv[2]);
    const int64_t W = std::atoll(argv[3]), H = std::atoll(argv[4]);
    std::vector<float> host(size_t(n * n * n));
    std::ifstream in(argv[1], std::ios::binary);
    in.read(reinterpret_cast<char*>(host.data()), std::streamsize(host.size() * sizeof(float)));
    if(!in)
    {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 2;
    }
    DensitySource src;
    src.X = n;
    src.Y = n;
    if(hipMalloc(reinterpret_cast<void**>(&src.ptr), host.size() * sizeof(float)) != hipSuccess
       || hipMemcpy(src.ptr, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    {
        std::fprintf(stderr, "device allocation failed\n");
        return 3;
    }
    std::tuple<DensitySource> volumeSources(src);
    std::tuple<> fieldSources;
    std::tuple<> particleSources;
    {
        isaac::IsaacVisualization<std::tuple<DensitySource>, std::tuple<>, std::tuple<>, 1024> vis(
            0, nullptr, "shim_example", 0, "localhost", 2460, {W, H}, {n, n, n}, {n, n, n}, {n / 4, n / 4, n / 4},
            {0, 0, 0}, volumeSources, fieldSources, particleSources, {1.0f, 1.0f, 1.0f});
        if(vis.init() != 0)
            return 4;
        vis.transferPoints[0].clear();
        vis.transferPoints[0][0] = {0.0f, 0.0f, 0.4f, 0.0f};
        vis.transferPoints[0][256] = {0.0f, 0.5f, 1.0f, 0.05f};
        vis.transferPoints[0][614] = {1.0f, 0.85f, 0.1f, 0.35f};
        vis.transferPoints[0][1024] = {1.0f, 0.15f, 0.1f, 0.9f};
        vis.updateTransfer();
        vis.interpolation = true;
        vis.step = 0.5f;
        for(float& c : vis.backgroundColor)
            c = 0.0f;
        vis.sendMinMax = true;
        // the reference's calling pattern (example/example.cpp:429-437,567; lib/isaac/mainpage.dox:113-117)
        json_object_set_new(vis.getJsonMetaRoot(), "kernel_time", json_string("kernel_time"));
        json_t* meta = vis.doVisualization(isaac::META_MASTER, nullptr, true);
        if(meta != nullptr)
            return 5;
        json_decref(meta);
        std::ofstream out(argv[5], std::ios::binary);
        out.write(reinterpret_cast<const char*>(vis.image().data()), std::streamsize(vis.image().size()));
        std::printf("shim_example: rendered %lldx%lld, kernel %llu us\n", (long long) W, (long long) H,
                    (unsigned long long) vis.kernelTime);
        std::printf("minmax %.9g %.9g\n", double(vis.minMaxArray.min[0]), double(vis.minMaxArray.max[0]));
        // the video frame message the reference's communicator would send (JPEG, quality 90)
        const std::string msg = vis.videoFrameMessage(90);
        std::ofstream mout(std::string(argv[5]) + ".msg", std::ios::binary);
        mout.write(msg.data(), std::streamsize(msg.size()));
        std::printf("video frame message %zu bytes\n", msg.size());
        // the server messages of this visualization: register (init), one periodic
        // message with every section, exit
        const std::string reg = vis.registerMessage(1);
        const std::string per = vis.periodMessage(0x1ffffffu);
        const std::string fin = decltype(vis)::exitMessage();
        for(const auto& kv : {std::make_pair(".register", &reg), std::make_pair(".period", &per), std::make_pair(".exit", &fin)})
        {
            std::ofstream f(std::string(argv[5]) + kv.first, std::ios::binary);
            f.write(kv.second->data(), std::streamsize(kv.second->size()));
        }
        std::printf("server messages %zu %zu %zu bytes\n", reg.size(), per.size(), fin.size());
    }
    (void) hipFree(src.ptr);
    return 0;
}
