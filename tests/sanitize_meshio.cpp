// The library's mesh reader (finitevolume_amd/csrc/meshio.cpp, host code)
// built alone with AddressSanitizer and UndefinedBehaviorSanitizer: reads
// each file named on the command line and prints its counts, or the
// reader's error.  Used by tests/test_sanitize.py on the committed meshes
// and on damaged copies of them.
#include <cstdarg>
#include <cstdio>
#include <string>

#include "../finitevolume_amd/csrc/meshio.cpp"

namespace fvk {
static std::string g_last;
int set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last = buf;
    return code;
}
}  // namespace fvk

int main(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
        fvk_mesh_file* h = nullptr;
        int rc = fvk_mesh_read(argv[i], &h);
        if (rc != FVK_OK) {
            std::printf("ERR %d %s\n", rc, fvk::g_last.c_str());
            continue;
        }
        fvk_mesh_desc d{};
        fvk_mesh_file_desc(h, &d);
        std::printf("OK %lld %lld %lld %d\n", (long long)d.n_nodes, (long long)d.n_cells, (long long)d.n_faces,
                    d.n_patches);
        fvk_mesh_file_destroy(h);
    }
    return 0;
}
