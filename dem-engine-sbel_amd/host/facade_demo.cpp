// A DEM-Engine program written against the DEMSolver API, built on libdeme. Used as the facade's end-to-end check
// and as the drop-in example.
//   facade_demo <hertz|frictionless|linear|prescribed> <n_spheres> <seconds>
//     Spheres dropped onto the floor of a box settle under gravity (the resting depth has a closed form for each
//     force model). prescribed: the Hertz model; sphere 0 in family 5 moves with the expression velocity (0.1 t, 0,
//     0); the other spheres start in family 0 and move to family 6 once they fall faster than 0.05 m/s
//     (ChangeFamilyWhen). Prints one JSON line: the model, the mean resting height of the sphere centres above the
//     floor, their largest speed, the number of contacts and the sphere radius, mass and material.
//   facade_demo io <n_clumps> <seconds> <dir>
//     Single spheres and three-sphere clumps in a box settle; then the sphere, clump and contact files are written
//     into <dir> with several content settings, with the raw state they were written from (state.json), and the
//     clump file is read back and used to start a second simulation, whose clump file is written too.
//   facade_demo inspect
//     Balls settle on the floor while two more fall from above; inspectors (whole domain and a region) and trackers
//     (of the falling pair and of an added plane) are read, one tracked ball is moved with SetPos, and the
//     inspectors are read again. Prints one JSON line.
//   facade_demo update <dir>
//     The frictionless model is read from a file (ReadContactForceModel); 18 balls settle on the floor; the cache is
//     cleared and three more balls are added above with UpdateClumps (tracked); one resting ball grows by 2%
//     (ChangeClumpSizes, by 2%); after the new balls land, the contact and sphere files are written into <dir>.
//   facade_demo readcsv <file>
//     Reads <file> with ReadClumpXyzFromCsv and ReadClumpQuatFromCsv and prints the values as float bit patterns.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "DEMSolver.h"

using namespace deme;

namespace {

uint32_t bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

int read_csv(const char* file) {
    std::string out = "{";
    try {
        auto xyz = DEMSolver::ReadClumpXyzFromCsv(file);
        auto q = DEMSolver::ReadClumpQuatFromCsv(file);
        std::map<std::string, std::vector<float3>> xs(xyz.begin(), xyz.end());
        std::map<std::string, std::vector<float4>> qs(q.begin(), q.end());
        out += "\"xyz\": {";
        bool first = true;
        for (const auto& kv : xs) {
            out += (first ? "\"" : ", \"") + kv.first + "\": [";
            first = false;
            for (size_t i = 0; i < kv.second.size(); i++) {
                const float3 v = kv.second[i];
                out += (i ? ", [" : "[") + std::to_string(bits(v.x)) + ", " + std::to_string(bits(v.y)) + ", " +
                       std::to_string(bits(v.z)) + "]";
            }
            out += "]";
        }
        out += "}, \"quat\": {";
        first = true;
        for (const auto& kv : qs) {
            out += (first ? "\"" : ", \"") + kv.first + "\": [";
            first = false;
            for (size_t i = 0; i < kv.second.size(); i++) {
                const float4 v = kv.second[i];
                out += (i ? ", [" : "[") + std::to_string(bits(v.w)) + ", " + std::to_string(bits(v.x)) + ", " +
                       std::to_string(bits(v.y)) + ", " + std::to_string(bits(v.z)) + "]";
            }
            out += "]";
        }
        out += "}}";
    } catch (const std::runtime_error& e) {
        // the message names the file; the test matches on the reason only
        std::string m = e.what();
        for (auto& ch : m)
            if (ch == '"' || ch == '\\') ch = '\'';
        out = "{\"error\": \"" + m + "\"}";
    }
    std::printf("%s\n", out.c_str());
    return 0;
}

// The io scene: n clumps alternating between a sphere (named "ball") and a three-sphere clump (left unnamed, so it
// takes the default name of its load order), stacked in columns above the floor of a box open at the top.
struct IoScene {
    std::shared_ptr<DEMClumpTemplate> ball, tri;
    std::shared_ptr<DEMMaterial> mat;
};

IoScene io_setup(DEMSolver& sim) {
    IoScene sc;
    sim.InstructBoxDomainDimension(0.2f, 0.2f, 0.2f);
    sim.SetCoordSysOrigin("center");
    sc.mat = sim.LoadMaterial({{"E", 1e7f}, {"nu", 0.3f}, {"CoR", 0.3f}, {"mu", 0.4f}, {"Crr", 0.f}});
    sim.InstructBoxDomainBoundingBC("top_open", sc.mat);
    const float rho = 2600.f, pi = 3.14159265f;
    const float rb = 0.004f, mb = rho * 4.f / 3.f * pi * rb * rb * rb;
    sc.ball = sim.LoadSphereType(mb, rb, sc.mat);
    sc.ball->AssignName("ball");
    const float rt = 0.003f, mt = 3.f * rho * 4.f / 3.f * pi * rt * rt * rt;
    const float It = 0.4f * mt * rt * rt + mt * 0.003f * 0.003f;
    sc.tri = sim.LoadClumpType(mt, make_float3(It, It, 1.5f * It), {rt, rt, rt},
                               {make_float3(-0.003f, 0, 0), make_float3(0.003f, 0, 0), make_float3(0, 0.003f, 0)},
                               sc.mat);
    sim.SetInitTimeStep(1e-5);
    sim.SetCDUpdateFreq(4);
    sim.SetMaxVelocity(2.f);
    sim.SetGravitationalAcceleration(make_float3(0, 0, -9.81f));
    return sc;
}

void dump_state(DEMSolver& sim, const IoScene& sc, const std::vector<std::shared_ptr<DEMClumpTemplate>>& types,
                const std::string& file) {
    deme_ctx* ctx = sim.GetContext();
    deme_counts cnt;
    deme_get_counts(ctx, &cnt);
    const size_t nO = cnt.nOwners, C = cnt.nContacts;
    std::vector<uint64_t> vox(nO);
    std::vector<uint16_t> lx(nO), ly(nO), lz(nO);
    std::vector<float> f[16];
    for (auto& v : f) v.resize(nO);
    std::vector<uint8_t> fam(nO);
    deme_owner_state st{vox.data(), lx.data(), ly.data(), lz.data(), f[0].data(), f[1].data(), f[2].data(),
                        f[3].data(), f[4].data(), f[5].data(), f[6].data(), f[7].data(), f[8].data(), f[9].data(),
                        f[10].data(), f[11].data(), f[12].data(), f[13].data(), f[14].data(), f[15].data(), fam.data()};
    if (deme_download_state(ctx, &st) != DEME_OK) throw std::runtime_error("deme_download_state failed");
    std::vector<uint32_t> idA(C), idB(C);
    std::vector<uint8_t> type(C);
    std::vector<float> F(3 * C), T(3 * C), cp(3 * C);
    if (C && (deme_download_contacts(ctx, idA.data(), idB.data(), type.data(), nullptr) != DEME_OK ||
              deme_download_forces(ctx, F.data(), T.data(), cp.data(), nullptr) != DEME_OK))
        throw std::runtime_error(std::string("contact download failed: ") + deme_last_error(ctx));
    const deme_sim_params& p = sim.GetSimParams();
    FILE* o = std::fopen(file.c_str(), "w");
    if (!o) throw std::runtime_error("cannot write " + file);
    std::fprintf(o, "{\"nvXp2\": %d, \"nvYp2\": %d, \"voxelSize\": %.17g, \"l\": %.17g, \"LBF\": [%.9g, %.9g, %.9g],\n",
                 p.nvXp2, p.nvYp2, p.voxelSize_exact, p.l_exact, (double)p.LBF[0], (double)p.LBF[1], (double)p.LBF[2]);
    auto ints = [&](const char* name, size_t n, auto get) {
        std::fprintf(o, "\"%s\": [", name);
        for (size_t i = 0; i < n; i++) std::fprintf(o, i ? ", %llu" : "%llu", (unsigned long long)get(i));
        std::fprintf(o, "],\n");
    };
    auto floats = [&](const char* name, size_t n, const float* v) {
        std::fprintf(o, "\"%s\": [", name);
        for (size_t i = 0; i < n; i++) std::fprintf(o, i ? ", %.9g" : "%.9g", (double)v[i]);
        std::fprintf(o, "],\n");
    };
    // voxel ids can exceed 2^53: written as strings
    std::fprintf(o, "\"voxel\": [");
    for (size_t i = 0; i < nO; i++) std::fprintf(o, i ? ", \"%llu\"" : "\"%llu\"", (unsigned long long)vox[i]);
    std::fprintf(o, "],\n");
    ints("locX", nO, [&](size_t i) { return lx[i]; });
    ints("locY", nO, [&](size_t i) { return ly[i]; });
    ints("locZ", nO, [&](size_t i) { return lz[i]; });
    const char* names[10] = {"qw", "qx", "qy", "qz", "vx", "vy", "vz", "wx", "wy", "wz"};
    for (int k = 0; k < 10; k++) floats(names[k], nO, f[k].data());
    ints("family", nO, [&](size_t i) { return fam[i]; });
    // component spheres in load order, and the template name of each clump
    std::vector<uint32_t> sOwner;
    std::vector<float> sR, sRel;
    std::fprintf(o, "\"clump_type\": [");
    for (size_t c = 0; c < types.size(); c++) {
        const auto& t = types[c];
        std::fprintf(o, c ? ", \"%s\"" : "\"%s\"", t->m_name.c_str());
        for (unsigned j = 0; j < t->nComp; j++) {
            sOwner.push_back((uint32_t)c);
            sR.push_back(t->radii[j]);
            sRel.push_back(t->relPos[j].x);
            sRel.push_back(t->relPos[j].y);
            sRel.push_back(t->relPos[j].z);
        }
    }
    std::fprintf(o, "],\n");
    ints("sph_owner", sOwner.size(), [&](size_t i) { return sOwner[i]; });
    floats("sph_r", sR.size(), sR.data());
    floats("sph_rel", sRel.size(), sRel.data());
    // the bounding box is one external object after the clumps: its five planes belong to owner n_clumps
    ints("anal_owner", 5, [&](size_t) { return types.size(); });
    ints("idA", C, [&](size_t i) { return idA[i]; });
    ints("idB", C, [&](size_t i) { return idB[i]; });
    ints("type", C, [&](size_t i) { return type[i]; });
    floats("force", 3 * C, F.data());
    floats("tof", 3 * C, T.data());
    floats("cpA", 3 * C, cp.data());
    std::fprintf(o, "\"n_clumps\": %zu}\n", types.size());
    std::fclose(o);
    (void)sc;
}

int run_io(int n, double seconds, const std::string& dir) {
    std::vector<std::shared_ptr<DEMClumpTemplate>> types;
    {
        DEMSolver sim;
        IoScene sc = io_setup(sim);
        std::vector<float3> xyz, vel;
        std::vector<float4> q;
        std::vector<unsigned int> fam;
        for (int i = 0; i < n; i++) {
            const int col = i % 16, layer = i / 16;
            xyz.push_back(make_float3(-0.024f + 0.016f * (col % 4), -0.024f + 0.016f * (col / 4),
                                      -0.1f + 0.0065f + 0.013f * layer));
            types.push_back(i % 2 ? sc.tri : sc.ball);
            const float a = 0.37f * i, s = std::sin(0.5f * a) / std::sqrt(1.3125f);
            float4 qi;
            qi.w = std::cos(0.5f * a);
            qi.x = s;
            qi.y = 0.5f * s;
            qi.z = 0.25f * s;
            q.push_back(qi);
            vel.push_back(make_float3(0.01f * std::sin((float)i), 0.01f * std::cos((float)i), 0.f));
            fam.push_back(i % 7 == 3 ? 1u : 0u);
        }
        auto batch = sim.AddClumps(types, xyz);
        batch->SetOriQ(q);
        batch->SetVel(vel);
        batch->SetFamilies(fam);
        sim.DisableFamilyOutput(1);
        sim.Initialize();
        sim.DoDynamics(seconds);
        sim.SetOutputContent(OUTPUT_CONTENT::QUAT | OUTPUT_CONTENT::ABSV | OUTPUT_CONTENT::VEL | OUTPUT_CONTENT::FAMILY);
        sim.WriteSphereFile(dir + "/spheres_all.csv");
        sim.WriteClumpFile(dir + "/clumps_all.csv");
        sim.SetContactOutputContent(CNT_OUTPUT_CONTENT::FORCE | CNT_OUTPUT_CONTENT::POINT | CNT_OUTPUT_CONTENT::NORMAL |
                                    CNT_OUTPUT_CONTENT::TORQUE_ONLY_FORCE);
        sim.WriteContactFile(dir + "/contacts_all.csv");
        sim.SetOutputContent(OUTPUT_CONTENT::QUAT | OUTPUT_CONTENT::ABSV);
        sim.WriteSphereFile(dir + "/spheres_default.csv");
        sim.WriteClumpFile(dir + "/clumps_default.csv");
        sim.SetContactOutputContent(CNT_OUTPUT_CONTENT::FORCE | CNT_OUTPUT_CONTENT::POINT);
        sim.WriteContactFile(dir + "/contacts_default.csv");
        sim.SetContactOutputContent(CNT_OUTPUT_CONTENT::OWNERS);
        sim.WriteContactFile(dir + "/contacts_owners.csv");
        dump_state(sim, sc, types, dir + "/state.json");
        std::printf("{\"contacts\": %zu, \"time\": %.12g}\n", sim.GetNumContacts(), sim.GetTimeElapsed());
    }
    // restart: a new simulation from the clump file (the clumps of family 1 were left out of it)
    auto xyz = DEMSolver::ReadClumpXyzFromCsv(dir + "/clumps_all.csv");
    auto q = DEMSolver::ReadClumpQuatFromCsv(dir + "/clumps_all.csv");
    DEMSolver sim2;
    IoScene sc2 = io_setup(sim2);
    for (const auto& kv : xyz) {
        auto t = kv.first == "ball" ? sc2.ball : sc2.tri;
        auto b = sim2.AddClumps(std::vector<std::shared_ptr<DEMClumpTemplate>>(kv.second.size(), t), kv.second);
        b->SetOriQ(q.at(kv.first));
    }
    sim2.Initialize();
    sim2.WriteClumpFile(dir + "/clumps_restart.csv");
    return 0;
}

int run_inspect() {
    DEMSolver sim;
    sim.InstructBoxDomainDimension(0.2f, 0.2f, 0.2f);
    sim.SetCoordSysOrigin("center");
    auto mat = sim.LoadMaterial({{"E", 1e7f}, {"nu", 0.3f}, {"CoR", 0.3f}, {"mu", 0.4f}, {"Crr", 0.f}});
    sim.InstructBoxDomainBoundingBC("top_open", mat);
    const float r = 0.004f, pi = 3.14159265f, vol = 4.f / 3.f * pi * r * r * r;
    auto ball = sim.LoadSphereType(2600.f * vol, r, mat);
    ball->SetVolume(vol);
    std::vector<float3> rest, fall = {make_float3(0.03f, 0.01f, 0.02f), make_float3(-0.05f, 0.02f, 0.04f)};
    for (int i = 0; i < 18; i++) rest.push_back(make_float3(-0.08f + 0.02f * (i % 9), -0.08f + 0.02f * (i / 9), -0.1f + r));
    auto b1 = sim.AddClumps(std::vector<std::shared_ptr<DEMClumpTemplate>>(rest.size(), ball), rest);
    auto b2 = sim.AddClumps(std::vector<std::shared_ptr<DEMClumpTemplate>>(fall.size(), ball), fall);
    auto wall = sim.AddBCPlane(make_float3(0, 0.09f, 0), make_float3(0, -1, 0), mat);
    wall->SetInitPos(make_float3(0.01f, 0.02f, 0.03f));
    auto tFall = sim.Track(b2);
    auto tWall = sim.Track(wall);
    sim.SetInitTimeStep(1e-5);
    sim.SetCDUpdateFreq(1);
    sim.SetGravitationalAcceleration(make_float3(0, 0, -9.81f));
    auto maxZ = sim.CreateInspector("clump_max_z");
    auto minZ = sim.CreateInspector("clump_min_z");
    auto maxV = sim.CreateInspector("clump_max_absv");
    auto volAll = sim.CreateInspector("clump_volume");
    auto volLeft = sim.CreateInspector("clump_volume", "return X < -0.01f;");
    auto maxZLow = sim.CreateInspector("clump_max_z", "return Z < -0.09f;");
    sim.Initialize();
    sim.DoDynamics(0.05);
    int nLeft = 0;
    double vmax = 0, zmax = -1;
    for (size_t o = 0; o < sim.GetNumClumps(); o++) {
        const float3 p = sim.GetOwnerPosition((uint32_t)o), v = sim.GetOwnerVelocity((uint32_t)o);
        nLeft += p.x < -0.01f;  // the grid has a column at x = 0: keep the boundary off it
        vmax = std::fmax(vmax, std::sqrt((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z));
        zmax = std::fmax(zmax, (double)p.z);
    }
    const float3 f0 = tFall->Pos(0), f1 = tFall->Pos(1), o18 = sim.GetOwnerPosition(18), w = tWall->Pos();
    const float before[6] = {maxZ->GetValue(), minZ->GetValue(), maxV->GetValue(), volAll->GetValue(),
                             volLeft->GetValue(), maxZLow->GetValue()};
    tFall->SetPos(make_float3(0.05f, 0.05f, 0.08f), 1);
    tFall->SetVel(make_float3(0.f, 0.f, 0.f), 1);
    const float after = maxZ->GetValue();
    std::string err;
    try {
        sim.CreateInspector("clump_volume", "X < 0")->GetValue();
    } catch (const std::runtime_error& e) {
        err = "region";
    }
    try {
        sim.CreateInspector("clump_max_y");
    } catch (const std::runtime_error& e) {
        err += ",quantity";
    }
    std::printf("{\"max_z\": %.9g, \"min_z\": %.9g, \"max_absv\": %.9g, \"volume\": %.9g, \"volume_left\": %.9g, "
                "\"max_z_low\": %.9g, \"max_z_after\": %.9g, \"vol\": %.9g, \"r\": %.9g, \"n\": %zu, "
                "\"n_left\": %d, \"vmax\": %.9g, \"zmax\": %.9g, \"fall0\": [%.9g, %.9g, %.9g], "
                "\"fall1\": [%.9g, %.9g, %.9g], \"owner18\": [%.9g, %.9g, %.9g], \"wall\": [%.9g, %.9g, %.9g], "
                "\"wall_owner\": %u, \"errors\": \"%s\"}\n",
                before[0], before[1], before[2], before[3], before[4], before[5], after, vol, r, sim.GetNumClumps(),
                nLeft, vmax, zmax, f0.x, f0.y, f0.z, f1.x, f1.y, f1.z, o18.x, o18.y, o18.z, w.x, w.y, w.z,
                tWall->obj->ownerID, err.c_str());
    return 0;
}

int run_update(const std::string& dir) {
    {
        std::FILE* f = std::fopen((dir + "/frictionless.model").c_str(), "w");
        if (!f) throw std::runtime_error("cannot write the model file");
        std::fputs(deme_builtin_force_model(1), f);
        std::fclose(f);
    }
    DEMSolver sim;
    sim.InstructBoxDomainDimension(0.2f, 0.2f, 0.2f);
    sim.SetCoordSysOrigin("center");
    auto mat = sim.LoadMaterial({{"E", 1e7f}, {"nu", 0.3f}, {"CoR", 0.3f}, {"mu", 0.4f}, {"Crr", 0.f}});
    sim.InstructBoxDomainBoundingBC("top_open", mat);
    auto fm = sim.ReadContactForceModel(dir + "/frictionless.model");
    fm->SetPerContactWildcards({});
    const float r = 0.004f, pi = 3.14159265f, vol = 4.f / 3.f * pi * r * r * r;
    auto ball = sim.LoadSphereType(2600.f * vol, r, mat);
    ball->AssignName("ball");
    std::vector<float3> rest;
    for (int i = 0; i < 18; i++) rest.push_back(make_float3(-0.08f + 0.02f * (i % 9), -0.08f + 0.02f * (i / 9), -0.1f + r));
    sim.AddClumps(std::vector<std::shared_ptr<DEMClumpTemplate>>(rest.size(), ball), rest);
    sim.SetInitTimeStep(1e-5);
    sim.SetCDUpdateFreq(1);
    sim.SetGravitationalAcceleration(make_float3(0, 0, -9.81f));
    sim.Initialize();
    sim.DoDynamics(0.02);
    sim.ClearCache();
    std::vector<float3> more = {make_float3(0.05f, 0.05f, -0.1f + r + 0.002f), make_float3(0.07f, 0.05f, -0.1f + r + 0.003f),
                                make_float3(0.05f, 0.07f, -0.1f + r + 0.004f)};
    auto b2 = sim.AddClumps(std::vector<std::shared_ptr<DEMClumpTemplate>>(more.size(), ball), more);
    auto t2 = sim.Track(b2);
    sim.UpdateClumps();
    const float3 start = t2->Pos(0);
    sim.ChangeClumpSizes({0}, {1.02f});
    sim.DoDynamics(0.05);
    sim.SetContactOutputContent(CNT_OUTPUT_CONTENT::FORCE);
    sim.WriteContactFile(dir + "/contacts.csv");
    sim.WriteSphereFile(dir + "/spheres.csv");
    const float3 p19 = sim.GetOwnerPosition(19), tp = t2->Pos(0), p0 = sim.GetOwnerPosition(0);
    std::printf("{\"n_clumps\": %zu, \"n_owners\": %zu, \"tracker_owner\": %u, \"span\": %zu, "
                "\"start\": [%.9g, %.9g, %.9g], \"tracker\": [%.9g, %.9g, %.9g], \"owner19\": [%.9g, %.9g, %.9g], "
                "\"owner0\": [%.9g, %.9g, %.9g], \"r\": %.9g}\n",
                sim.GetNumClumps(), sim.GetNumOwners(), t2->obj->ownerID, t2->obj->nSpanOwners, start.x, start.y,
                start.z, tp.x, tp.y, tp.z, p19.x, p19.y, p19.z, p0.x, p0.y, p0.z, r);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc == 3 && std::string(argv[1]) == "readcsv") return read_csv(argv[2]);
    if (argc == 3 && std::string(argv[1]) == "update") {
        try {
            return run_update(argv[2]);
        } catch (const std::exception& e) {
            std::fprintf(stderr, "facade_demo: %s\n", e.what());
            return 1;
        }
    }
    if (argc == 2 && std::string(argv[1]) == "inspect") {
        try {
            return run_inspect();
        } catch (const std::exception& e) {
            std::fprintf(stderr, "facade_demo: %s\n", e.what());
            return 1;
        }
    }
    if (argc == 5 && std::string(argv[1]) == "io") {
        try {
            return run_io(std::atoi(argv[2]), std::atof(argv[3]), argv[4]);
        } catch (const std::exception& e) {
            std::fprintf(stderr, "facade_demo: %s\n", e.what());
            return 1;
        }
    }
    if (argc != 4) {
        std::fprintf(stderr, "usage: facade_demo <hertz|frictionless|linear|prescribed> <n_spheres> <seconds> | "
                             "io <n_clumps> <seconds> <dir> | inspect | update <dir> | readcsv <file>\n");
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
