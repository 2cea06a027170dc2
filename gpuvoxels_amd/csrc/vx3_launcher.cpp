// Voxelyze3 -- runs one generation directory through a worker and copies its report.
//
//   Voxelyze3 -i <data_path> [-o <report>] [-w <worker>] [-l] [-f]
//
// Command-line surface of Voxelyze3/src/Executables/Voxelyze3.cpp:25-128, the entry
// point the reference's experiment scripts call
// (Experiments/framework/voxelyze/workflow.py:111-117): the data path holds base.vxa
// and *.vxd files (extension matched case-insensitively); a task file listing them is
// written to workspace/locally/<YYYYmmddHHMMSS>.vxt under the current directory, the
// worker runs as `<worker> -i <vxt> -o <vxr>`, and the report is copied to the output
// path.  Exit status 1 on bad arguments, an existing output without -f, a missing data
// path, base.vxa or worker; 0 otherwise (also when the worker fails, as there).
// Differences: VXD files are listed in name order (the reference lists them in
// directory order); without -w, ./vx3_node_worker is used if present, else the worker
// next to this executable; -n (several nodes, not implemented in the reference
// either) only prints that it is not available.
#include <dirent.h>
#include <sys/stat.h>
#include <sys/wait.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <string>
#include <vector>

namespace {

const char* USAGE =
    "Thank you for using Voxelyze3. This program should be run on a computer that has GPUs.\n"
    "Typical Usage:\n"
    "Voxelyze -i <data_path> -o <report_path> -l -f \n\n"
    "Allowed options:\n"
    "  -h [ --help ]          produce help message\n"
    "  -l [ --locally ]       If this machine already has GPUs, locally run tasks on this machine.\n"
    "  -n [ --nodes ] arg     The number of nodes you want to start simultaneously. (Not implemented.)\n"
    "  -i [ --input ] arg     Set input directory path which contains a generation of VXA files.\n"
    "  -o [ --output ] arg    Set output file path for report. (e.g. report_1.xml)\n"
    "  -w [ --worker ] arg    Specify which worker you want to use. vx3_node_worker by default.\n"
    "  -f [ --force ]         Overwrite output file if exists.\n";

bool isFile(const std::string& p) {
    struct stat s;
    return stat(p.c_str(), &s) == 0 && S_ISREG(s.st_mode);
}

bool isDir(const std::string& p) {
    struct stat s;
    return stat(p.c_str(), &s) == 0 && S_ISDIR(s.st_mode);
}

std::string join(const std::string& a, const std::string& b) {
    if (a.empty()) return b;
    return a.back() == '/' ? a + b : a + "/" + b;
}

std::string lower(std::string s) {
    for (char& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

// text escaped for an XML element body
std::string xmlText(const std::string& s) {
    std::string o;
    for (char c : s) {
        switch (c) {
        case '&': o += "&amp;"; break;
        case '<': o += "&lt;"; break;
        case '>': o += "&gt;"; break;
        default: o += c;
        }
    }
    return o;
}

std::string selfDir() {
    char buf[4096];
    ssize_t n = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
    if (n <= 0) return ".";
    buf[n] = 0;
    std::string p(buf);
    size_t k = p.find_last_of('/');
    return k == std::string::npos ? "." : p.substr(0, k);
}

bool copyFile(const std::string& from, const std::string& to) {
    std::ifstream in(from, std::ios::binary);
    std::ofstream out(to, std::ios::binary | std::ios::trunc);
    if (!in || !out) return false;
    out << in.rdbuf();
    return (bool)out;
}

}  // namespace

int main(int argc, char** argv) {
    std::string input, output, worker;
    bool help = false, locally = false, force = false, nodes = false, bad = false;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto value = [&](std::string& dst) {
            if (i + 1 < argc) dst = argv[++i];
            else bad = true;
        };
        std::string ignored;
        if (a == "--help") help = true;
        else if (a == "--locally") locally = true;
        else if (a == "--force") force = true;
        else if (a == "--input") value(input);
        else if (a == "--output") value(output);
        else if (a == "--worker") value(worker);
        else if (a == "--nodes") { nodes = true; value(ignored); }
        else if (a.size() >= 2 && a[0] == '-' && a[1] != '-') {
            // grouped short flags (-lf); a value option takes the rest of the word or the next one
            for (size_t k = 1; k < a.size(); k++) {
                const char c = a[k];
                std::string* dst = c == 'i' ? &input : c == 'o' ? &output : c == 'w' ? &worker : c == 'n' ? &ignored : nullptr;
                if (c == 'n') nodes = true;
                if (dst) {
                    if (k + 1 < a.size()) *dst = a.substr(k + 1);
                    else value(*dst);
                    break;
                }
                if (c == 'h') help = true;
                else if (c == 'l') locally = true;
                else if (c == 'f') force = true;
                else bad = true;
            }
        } else bad = true;
    }
    if (bad || help || input.empty()) { printf("%s\n", USAGE); return 1; }
    const bool hasOutput = !output.empty();
    if (nodes && locally) {
        printf("ERROR: -n and -s cannot use together.\n%s\n", USAGE);
        return 1;
    }
    if (!nodes) printf("Running simulation locally by default.\n");
    if (hasOutput && isFile(output) && !force) {
        printf("Error: output file exists.\n\n%s\n", USAGE);
        return 1;
    }
    if (!isDir(input)) {
        printf("Error: input directory not found.\n\n%s\n", USAGE);
        return 1;
    }
    if (!isFile(join(input, "base.vxa"))) {
        printf("No base.vxa found in input directory.\n\n%s\n", USAGE);
        return 1;
    }
    if (worker.empty()) {
        worker = "./vx3_node_worker";
        if (!isFile(worker) && isFile(join(selfDir(), "vx3_node_worker"))) worker = join(selfDir(), "vx3_node_worker");
    }
    if (!isFile(worker)) {
        printf("Need an executable worker but nothing found.\n\n%s\n", USAGE);
        return 1;
    }
    mkdir("workspace", 0755);
    if (nodes) {
        fprintf(stderr, "Running on several nodes is not available; nothing was run.\n");
        return 0;
    }
    const std::string locallyDir = "workspace/locally";
    mkdir(locallyDir.c_str(), 0755);

    std::vector<std::string> vxds;
    if (DIR* d = opendir(input.c_str())) {
        while (dirent* e = readdir(d)) {
            std::string name = e->d_name;
            size_t dot = name.find_last_of('.');
            if (dot == std::string::npos || dot == 0 || lower(name.substr(dot)) != ".vxd") continue;
            if (isFile(join(input, name))) vxds.push_back(name);
        }
        closedir(d);
    }
    std::sort(vxds.begin(), vxds.end());

    char stamp[32];
    time_t now = time(nullptr);
    strftime(stamp, sizeof(stamp), "%Y%m%d%H%M%S", localtime(&now));
    const std::string vxt = join(locallyDir, std::string(stamp) + ".vxt");
    const std::string vxr = join(locallyDir, std::string(stamp) + ".vxr");
    {
        std::ofstream f(vxt, std::ios::binary | std::ios::trunc);
        f << "<?xml version=\"1.0\" encoding=\"utf-8\"?>\n<vxa>" << xmlText(join(input, "base.vxa")) << "</vxa><input_dir>"
          << xmlText(input) << "</input_dir>";
        if (!vxds.empty()) {
            f << "<vxd>";
            for (const std::string& n : vxds) f << "<f>" << xmlText(n) << "</f>";
            f << "</vxd>";
        }
        f << "\n";
        if (!f) {
            printf("ERROR: cannot write %s.\n", vxt.c_str());
            return 1;
        }
    }
    printf("%s -i %s -o %s\n", worker.c_str(), vxt.c_str(), vxr.c_str());
    fflush(stdout);
    pid_t pid = fork();
    if (pid == 0) {
        execl(worker.c_str(), worker.c_str(), "-i", vxt.c_str(), "-o", vxr.c_str(), (char*)nullptr);
        _exit(127);
    }
    if (pid > 0) {
        int st = 0;
        while (waitpid(pid, &st, 0) < 0 && errno == EINTR) {}
    }
    if (isFile(vxr)) {
        if (hasOutput && !copyFile(vxr, output)) printf("ERROR: Failed to copy result file: %s.\n", vxr.c_str());
    } else {
        printf("File not exist: %s. Worker failed to finish the job.\n", vxr.c_str());
    }
    return 0;
}
