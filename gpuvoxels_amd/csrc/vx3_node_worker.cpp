// vx3_node_worker -- runs the simulations of a task file and writes their report.
//
//   vx3_node_worker -i task.vxt -o report.xml [-f] [--f32] [--allow-unsupported] [--voxcad-stop]
//
// Command-line surface of Voxelyze3/src/Executables/vx3_node_worker.cu:32-143: the
// task file holds <vxa>base.vxa</vxa>, <input_dir>dir</input_dir> and
// <vxd><f>name.vxd</f>...</vxd> (as Voxelyze3.cpp:101-111 writes it); VXD files are
// read from input_dir and spread over the devices round-robin; the report is written
// to the output path.  Exit status 1 on bad arguments, an existing output without -f,
// a missing task file or no GPU; 0 otherwise.  Additions: --f32 runs the FP32 state
// engine (default FP64), --allow-unsupported is accepted for scripts that pass it
// (every feature the loader reads is simulated), and --voxcad-stop honours the VXA's StopConditionType/Value and EquilibriumMode the way
// VoxCAD does (the Voxelyze3 worker reads and ignores them).
#include <sys/stat.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/vx3h.h"
#include "vx3_manager.h"

namespace {

const char* USAGE =
    "Usage: vx3_node_worker -i task.vxt -o report.xml [-f]\n"
    "Allowed options:\n"
    "  -h [ --help ]          produce help message\n"
    "  -i [ --input ] arg     task file (vxa, input_dir, vxd list)\n"
    "  -o [ --output ] arg    output file path for the report (e.g. report_1.xml)\n"
    "  -f [ --force ]         overwrite output file if exists\n"
    "  --f32                  FP32 state engine (default FP64)\n"
    "  --allow-unsupported    accepted; every feature the loader reads is simulated\n"
    "  --voxcad-stop          honour StopConditionType/Value and EquilibriumMode as VoxCAD does\n";

bool isFile(const std::string& p) {
    struct stat s;
    return stat(p.c_str(), &s) == 0 && S_ISREG(s.st_mode);
}

std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}

// the five predefined XML entities decoded
std::string unescape(const std::string& s) {
    static const char* ent[5][2] = {{"&amp;", "&"}, {"&lt;", "<"}, {"&gt;", ">"}, {"&quot;", "\""}, {"&apos;", "'"}};
    std::string o;
    for (size_t i = 0; i < s.size();) {
        bool hit = false;
        if (s[i] == '&')
            for (auto& e : ent)
                if (s.compare(i, strlen(e[0]), e[0]) == 0) { o += e[1]; i += strlen(e[0]); hit = true; break; }
        if (!hit) o += s[i++];
    }
    return o;
}

// text of every <tag>...</tag> in s (flat; the task file has no nesting beyond <vxd><f>)
std::vector<std::string> elements(const std::string& s, const std::string& tag, bool decode = true) {
    std::vector<std::string> out;
    const std::string open = "<" + tag + ">", close = "</" + tag + ">";
    size_t p = 0;
    while ((p = s.find(open, p)) != std::string::npos) {
        size_t q = s.find(close, p + open.size());
        if (q == std::string::npos) break;
        const std::string body = trim(s.substr(p + open.size(), q - p - open.size()));
        out.push_back(decode ? unescape(body) : body);
        p = q + close.size();
    }
    return out;
}

std::string dirName(const std::string& p) {
    std::string s = p;
    while (s.size() > 1 && s.back() == '/') s.pop_back();
    size_t k = s.find_last_of('/');
    return k == std::string::npos ? s : s.substr(k + 1);
}

}  // namespace

int main(int argc, char** argv) {
    std::string input, output;
    bool force = false, help = false;
    vx3::ManagerOptions opt;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto value = [&](std::string& dst) { if (i + 1 < argc) dst = argv[++i]; else help = true; };
        if (a == "-h" || a == "--help") help = true;
        else if (a == "-i" || a == "--input") value(input);
        else if (a == "-o" || a == "--output") value(output);
        else if (a == "-f" || a == "--force") force = true;
        else if (a == "--f32") opt.precision = VX3H_PREC_F32;
        else if (a == "--allow-unsupported") opt.allow_unsupported = true;
        else if (a == "--voxcad-stop") opt.voxcad_stop = true;
        else help = true;
    }
    if (help || input.empty() || output.empty()) { printf("%s\n", USAGE); return 1; }
    if (isFile(output) && !force) { printf("Error: output file exists.\n\n%s\n", USAGE); return 1; }
    if (!isFile(input)) { printf("Error: input file not found.\n\n%s\n", USAGE); return 1; }

    std::ifstream tf(input);
    std::stringstream ss;
    ss << tf.rdbuf();
    const std::string task = ss.str();
    std::vector<std::string> vxa = elements(task, "vxa"), dir = elements(task, "input_dir");
    std::vector<std::string> vxdBlock = elements(task, "vxd", false);
    std::vector<std::string> files;
    for (const std::string& blk : vxdBlock) for (const std::string& f : elements(blk, "f")) files.push_back(f);
    if (vxa.empty() || dir.empty()) { printf("Error: task file lacks <vxa> or <input_dir>.\n"); return 1; }

    int nDevices = 0;
    if (vx3h_device_count(&nDevices) != VX3H_OK || nDevices <= 0) { printf("ERROR: No GPU found.\n"); return 1; }
    printf("%d GPU found.\n", nDevices);
    std::vector<std::vector<std::string>> sub(nDevices);
    for (size_t i = 0; i < files.size(); i++) sub[i % nDevices].push_back(files[i]);

    vx3::SimulationManager sm(sub, vxa[0], dir[0], nDevices, opt);
    try {
        sm.start();
    } catch (const std::exception& e) {
        fprintf(stderr, "ERROR: %s\n", e.what());
        return 1;
    }
    std::vector<vx3h_report_entry> entries;
    for (const vx3::SimulationResult& r : sm.h_results) {
        vx3h_report_entry e{};
        e.vxd_name = r.vxa_filename.c_str();
        e.status = &r.status;
        e.voxel_size = r.voxSize;
        e.save_positions = r.SavePositionOfAllVoxels;
        e.init_pos = r.voxel_init_pos.data();
        e.pos = r.voxel_position.data();
        e.mats = r.voxel_mats.data();
        entries.push_back(e);
    }
    size_t len = 0;
    if (vx3h_report_xml(dirName(dir[0]).c_str(), entries.data(), (int)entries.size(), nullptr, 0, &len) != VX3H_OK) {
        fprintf(stderr, "ERROR: %s\n", vx3h_last_error());
        return 1;
    }
    std::vector<char> buf(len);
    if (vx3h_report_xml(dirName(dir[0]).c_str(), entries.data(), (int)entries.size(), buf.data(), len, &len) != VX3H_OK) {
        fprintf(stderr, "ERROR: %s\n", vx3h_last_error());
        return 1;
    }
    std::ofstream out(output, std::ios::binary);
    out << buf.data();
    return out ? 0 : 1;
}
