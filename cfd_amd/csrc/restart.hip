// Restart files of the reference (RESTART, ns2DComp.ALE.f90:394-435, and
// PRINTREST, :898-917), read and written on the host.
//
// Format: a Fortran unformatted sequential file, every record framed by its
// length in bytes as a 4-byte integer before and after the payload (little
// endian, as the reference's compiler writes it):
//   record 0        ITER (int32), TIME (binary64)                 12 bytes
//   record 1..npoin U(1:4,i), T(i), GAMM(i) (binary64)             48 bytes
// RESTART reads the first npoin node records and ignores what follows.
//
// The session entries checkpoint and restore a device-resident session
// through its field accessors (the caller's numbering on either side).
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "cfd_internal.hpp"

using namespace cfd;

namespace {

struct File {
    std::FILE* f = nullptr;
    explicit File(std::FILE* p) : f(p) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

void put_record(std::FILE* f, const void* payload, uint32_t n, const std::string& path) {
    if (std::fwrite(&n, 4, 1, f) != 1 || (n && std::fwrite(payload, 1, n, f) != n) ||
        std::fwrite(&n, 4, 1, f) != 1)
        throw Error("arg: restart file: write failed: " + path);
}

// one record of exactly n payload bytes
void get_record(std::FILE* f, void* payload, uint32_t n, const std::string& path, long long rec) {
    uint32_t a = 0, b = 0;
    if (std::fread(&a, 4, 1, f) != 1)
        throw Error("arg: restart file: ends before record " + std::to_string(rec) + ": " + path);
    if (a != n)
        throw Error("arg: restart file: record " + std::to_string(rec) + " holds " +
                    std::to_string(a) + " bytes, expected " + std::to_string(n) + ": " + path);
    if (std::fread(payload, 1, n, f) != n || std::fread(&b, 4, 1, f) != 1 || b != a)
        throw Error("arg: restart file: record " + std::to_string(rec) + " is truncated: " + path);
}

void write_restart(const char* path, int iter, double time, int npoin, const double* U,
                   const double* T, const double* GAMM) {
    CFD_ARG(path && npoin > 0 && U && T && GAMM, "null argument or npoin <= 0");
    File fh(std::fopen(path, "wb"));
    CFD_CHECK(fh.f, std::string("arg: restart file: cannot open for writing: ") + path);
    unsigned char head[12];
    std::memcpy(head, &iter, 4);
    std::memcpy(head + 4, &time, 8);
    put_record(fh.f, head, 12, path);
    for (int i = 0; i < npoin; ++i) {
        const double rec[6] = {U[4 * (size_t)i], U[4 * (size_t)i + 1], U[4 * (size_t)i + 2],
                               U[4 * (size_t)i + 3], T[i], GAMM[i]};
        put_record(fh.f, rec, 48, path);
    }
    CFD_CHECK(std::fflush(fh.f) == 0, std::string("arg: restart file: write failed: ") + path);
}

void read_restart(const char* path, int npoin, int* iter, double* time, double* U, double* T,
                  double* GAMM) {
    CFD_ARG(path && npoin > 0 && U && T && GAMM, "null argument or npoin <= 0");
    File fh(std::fopen(path, "rb"));
    CFD_CHECK(fh.f, std::string("arg: restart file: cannot open: ") + path);
    unsigned char head[12];
    get_record(fh.f, head, 12, path, 0);
    int it = 0;
    double t = 0.0;
    std::memcpy(&it, head, 4);
    std::memcpy(&t, head + 4, 8);
    for (int i = 0; i < npoin; ++i) {
        double rec[6];
        get_record(fh.f, rec, 48, path, i + 1);
        for (int q = 0; q < 4; ++q) U[4 * (size_t)i + q] = rec[q];
        T[i] = rec[4];
        GAMM[i] = rec[5];
    }
    if (iter) *iter = it;
    if (time) *time = t;
}

}  // namespace

extern "C" {

int cfd_write_restart(const char* path, int iter, double time, int npoin, const double* U,
                      const double* T, const double* GAMM) {
    CFD_TRY({ write_restart(path, iter, time, npoin, U, T, GAMM); })
}

int cfd_read_restart(const char* path, int npoin, int* iter, double* time, double* U, double* T,
                     double* GAMM) {
    CFD_TRY({ read_restart(path, npoin, iter, time, U, T, GAMM); })
}

int cfd_session_write_restart(cfd_session* S, const char* path) {
    CFD_TRY({
        CFD_ARG(S && path, "null argument");
        long long n = 0;
        CFD_CHECK(cfd_session_field_size(S, CFD_F_T, &n) == CFD_OK, cfd_last_error());
        std::vector<double> U(4 * (size_t)n), T(n), G(n), sc(CFD_NSCALARS);
        for (auto [f, p] : {std::pair<int, double*>{CFD_F_U, U.data()}, {CFD_F_T, T.data()},
                            {CFD_F_GAMM, G.data()}})
            CFD_CHECK(cfd_session_get_field(S, f, p) == CFD_OK, cfd_last_error());
        CFD_CHECK(cfd_session_get_scalars(S, sc.data()) == CFD_OK, cfd_last_error());
        write_restart(path, (int)sc[CFD_S_ITER], sc[CFD_S_TIME], (int)n, U.data(), T.data(),
                      G.data());
    })
}

int cfd_session_read_restart(cfd_session* S, const char* path, int keep_clock) {
    CFD_TRY({
        CFD_ARG(S && path && (keep_clock == 0 || keep_clock == 1), "null argument or keep_clock");
        long long n = 0;
        CFD_CHECK(cfd_session_field_size(S, CFD_F_T, &n) == CFD_OK, cfd_last_error());
        std::vector<double> U(4 * (size_t)n), T(n), G(n), VX(n), VY(n), RHO(n), E(n);
        int iter = 0;
        double time = 0.0;
        read_restart(path, (int)n, &iter, &time, U.data(), T.data(), G.data());
        // ns2DComp.ALE.f90:78: with NGAS /= 1 the loaded GAMM is replaced by
        // GAMA, which the session's GAMM already holds: GAMM is not loaded
        for (long long i = 0; i < n; ++i) {
            const double rho = U[4 * i];
            RHO[i] = rho;
            VX[i] = U[4 * i + 1] / rho;
            VY[i] = U[4 * i + 2] / rho;
            E[i] = U[4 * i + 3] / rho;
        }
        for (auto [f, p] : {std::pair<int, double*>{CFD_F_U, U.data()}, {CFD_F_U1, U.data()},
                            {CFD_F_T, T.data()}, {CFD_F_VEL_X, VX.data()},
                            {CFD_F_VEL_Y, VY.data()}, {CFD_F_RHO, RHO.data()}, {CFD_F_E, E.data()}})
            CFD_CHECK(cfd_session_set_field(S, f, p) == CFD_OK, cfd_last_error());
        std::vector<double> sc(CFD_NSCALARS);
        CFD_CHECK(cfd_session_get_scalars(S, sc.data()) == CFD_OK, cfd_last_error());
        // the main program restarts its clock and step control after RESTART
        // (ns2DComp.ALE.f90:107-110, 134); keep_clock = 1 keeps ITER and TIME
        sc[CFD_S_TIME] = keep_clock ? time : 0.0;
        sc[CFD_S_ITER] = keep_clock ? (double)iter : 0.0;
        sc[CFD_S_DTMIN] = 0.0;
        sc[CFD_S_DTMIN1] = 0.0;
        sc[CFD_S_BANDERA] = 1.0;
        CFD_CHECK(cfd_session_set_scalars(S, sc.data()) == CFD_OK, cfd_last_error());
    })
}

}  // extern "C"
