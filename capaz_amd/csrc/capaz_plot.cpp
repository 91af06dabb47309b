// Tecplot plot rows (include/capaz.h: capaz_write_plot_rows): the host-side
// printf loop of print_plot_row (main.cpp:700-780), so that plots of large
// grids are written at C speed and with the C library's number formatting,
// the reference's own.
#include <cstdio>
#include <cstring>
#include <string>

#include "capaz_ctx.h"

int capaz_write_plot_rows(const char* path, const double* table, long nrows, int ncols,
                          const char* const* formats) {
    using capaz::detail::fail;
    if (!path || nrows < 0 || ncols <= 0 || (nrows > 0 && !table) || !formats)
        return fail(nullptr, CAPAZ_ERR_INVALID, "capaz_write_plot_rows: bad argument");
    for (int c = 0; c < ncols; ++c) {
        // one double conversion per column, nothing else (the reference's
        // formats are %.2e, %.3e, %.5e)
        const char* f = formats[c];
        if (!f || f[0] != '%' || std::strchr(f + 1, '%') || !std::strpbrk(f, "eEfFgG") ||
            std::strpbrk(f, "sSnpcdiuxXo*"))
            return fail(nullptr, CAPAZ_ERR_INVALID, "capaz_write_plot_rows: each format must convert one double");
    }
    FILE* fp = std::fopen(path, "a");
    if (!fp) return fail(nullptr, CAPAZ_ERR_INVALID, std::string("capaz_write_plot_rows: cannot open ") + path);
    for (long r = 0; r < nrows; ++r) {
        const double* row = table + (size_t)r * (size_t)ncols;
        for (int c = 0; c < ncols; ++c) {
            if (c) std::fputc(' ', fp);
            std::fprintf(fp, formats[c], row[c]);
        }
        std::fputc('\n', fp);
    }
    const bool bad = std::ferror(fp) != 0;
    if (std::fclose(fp) != 0 || bad)
        return fail(nullptr, CAPAZ_ERR_INVALID, std::string("capaz_write_plot_rows: write to ") + path + " failed");
    return CAPAZ_OK;
}
