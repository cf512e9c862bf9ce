// A DEM-Engine program written against the DEMSolver API, built on libdeme: spheres dropped onto the floor of a
// box settle under gravity. Used as the facade's end-to-end check (the resting depth has a closed form for each
// force model) and as the drop-in example.
//   facade_demo <hertz|frictionless|linear|prescribed> <n_spheres> <seconds>
// prescribed: the Hertz model; sphere 0 in family 5 moves with the expression velocity (0.1 t, 0, 0); the other
// spheres start in family 0 and move to family 6 once they fall faster than 0.05 m/s (ChangeFamilyWhen).
// Prints one JSON line: the model, the mean resting height of the sphere centres above the floor, their largest
// speed, the number of contacts and the sphere radius, mass and material.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "DEMSolver.h"

using namespace deme;

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: facade_demo <hertz|frictionless|linear> <n_spheres> <seconds>\n");
        return 2;
    }
    const std::string model = argv[1];
    const int n = std::atoi(argv[2]);
    const double seconds = std::atof(argv[3]);
    const float r = 0.004f, rho = 2600.f;
    const float mass = rho * 4.f / 3.f * 3.14159265f * r * r * r;
    const float E = 1e7f, nu = 0.3f, CoR = 0.3f, mu = 0.4f;
    const float k = 2e4f;  // linear model stiffness, N/m
    try {
        DEMSolver sim;
        sim.InstructBoxDomainDimension(0.2f, 0.2f, 0.2f);
        sim.SetCoordSysOrigin("center");  // the box spans [-0.1, 0.1] on each axis
        auto mat = sim.LoadMaterial({{"E", E}, {"nu", nu}, {"CoR", CoR}, {"mu", mu}, {"Crr", 0.f}});
        sim.InstructBoxDomainBoundingBC("top_open", mat);
        if (model == "frictionless") {
            sim.UseFrictionlessHertzianModel();
        } else if (model == "linear") {
            // linear spring with a dashpot on the normal relative velocity, in the reference's model-string names
            auto fm = sim.DefineContactForceModel(
                "const float vn = dot(ALinVel - BLinVel, B2A);\n"
                "force = (float)(" + std::to_string(k) + " * overlapDepth) * B2A - (3.0f * vn) * B2A;\n");
            fm->SetPerContactWildcards({});
        } else if (model == "prescribed") {
            sim.SetFamilyPrescribedLinVel(5, "0.1 * t", "0", "0");
            sim.ChangeFamilyWhen(0, 6, "return vel.z < -0.05f;");
        } else if (model != "hertz") {
            std::fprintf(stderr, "unknown model %s\n", model.c_str());
            return 2;
        }
        auto tpl = sim.LoadSphereType(mass, r, mat);
        // a row of spheres 1 mm above the floor (the box floor is at z = -0.1), spaced well apart
        std::vector<float3> xyz;
        for (int i = 0; i < n; i++) xyz.push_back(make_float3(-0.08f + 0.02f * (i % 9), -0.08f + 0.02f * (i / 9), -0.1f + r + 0.001f));
        auto batch = sim.AddClumps(std::vector<std::shared_ptr<DEMClumpTemplate>>(xyz.size(), tpl), xyz);
        if (model == "prescribed") {
            std::vector<unsigned int> fam(xyz.size(), 0);
            fam[0] = 5;
            batch->SetFamilies(fam);
        }
        sim.SetInitTimeStep(1e-5);
        sim.SetCDUpdateFreq(1);
        sim.SetGravitationalAcceleration(make_float3(0, 0, -9.81f));
        sim.Initialize();
        sim.DoDynamics(seconds);
        double hsum = 0, vmax = 0;
        for (int i = 0; i < n; i++) {
            const float3 p = sim.GetOwnerPosition(i);
            const float3 v = sim.GetOwnerVelocity(i);
            hsum += (double)p.z + 0.1;
            vmax = std::fmax(vmax, std::sqrt((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z));
        }
        std::string fams;
        for (int i = 0; i < n; i++) fams += (i ? "," : "") + std::to_string(sim.GetOwnerFamily(i));
        const float3 v0 = sim.GetOwnerVelocity(0);
        std::printf("{\"model\": \"%s\", \"height\": %.9g, \"vmax\": %.6g, \"contacts\": %zu, \"r\": %.9g, \"mass\": %.9g, "
                    "\"E\": %g, \"nu\": %g, \"k\": %g, \"time\": %.12g, \"v0\": [%.9g, %.9g, %.9g], \"families\": [%s]}\n",
                    model.c_str(), hsum / n, vmax, sim.GetNumContacts(), (double)r, (double)mass, (double)E, (double)nu,
                    (double)k, sim.GetTimeElapsed(), (double)v0.x, (double)v0.y, (double)v0.z, fams.c_str());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "facade_demo: %s\n", e.what());
        return 1;
    }
    return 0;
}
