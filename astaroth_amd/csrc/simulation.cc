// Snapshot files, the diagnostics time series and the time-stepping driver of the reference's
// standalone simulation (samples/standalone/simulation.cc), restated over this library's Device:
//
//   mesh_info.list         simulation.cc:62-110   every parameter by name, with AcReal size and
//                                                 byte order, one per line
//   <FIELD>_<step>.mesh    simulation.cc:115-147  the time stamp, then the whole m^3 buffer
//                          (read: 205-245)        (ghosts included), AcReal, x fastest
//   timeseries.ts          simulation.cc:251-313  title row, then one row per diagnostics step:
//                          (title: 345-362)       step, t, dt, min/rms/max of |u| and of every
//                                                 field
//   driver loop            simulation.cc:316-560  CFL dt, step, diagnostics every AC_save_steps,
//                                                 snapshot every AC_bin_steps or AC_bin_save_t,
//                                                 stop at AC_max_time / small dt / NaN / STOP file
//
// The reference's analysis package reads these files (analysis/python/astar/data/read.py:33-66,
// 612-652); astaroth_amd/snapshot.py reads them the same way. Differences from the reference:
// the boundary fill is the corrected general fill (acDeviceGeneralBoundconds; periodic unless
// the configuration says otherwise), the diagnostics are FP64-accumulated device reductions,
// file errors return AC_FAILURE instead of printing and continuing, and a snapshot read checks
// the file length.
#include <cmath>
#include <cstdio>
#include <string>
#include <unistd.h>

#include "astaroth.h"
#include "astaroth_ext.h"

namespace {

std::string
in_dir(const char* dir, const std::string& name)
{
    if (!dir || !*dir)
        return name;
    std::string d(dir);
    if (d.back() != '/')
        d += '/';
    return d + name;
}

std::string
mesh_file(const char* dir, int field, int step)
{
    return in_dir(dir, std::string(vtxbuf_names[field]) + "_" + std::to_string(step) + ".mesh");
}

int3
mk3(int x, int y, int z)
{
    int3 v;
    v.x = x;
    v.y = y;
    v.z = z;
    return v;
}

// min, rms (sqrt(sum / n), the Node-level finalisation of node.cc:932-957) and max
struct Stats {
    double min, rms, max;
};

bool
field_stats(Device d, Stream s, double ncells, int f, Stats* out)
{
    double sum = 0;
    if (acDeviceReduceScalDouble(d, s, RTYPE_MIN, (VertexBufferHandle)f, &out->min) != AC_SUCCESS ||
        acDeviceReduceScalDouble(d, s, RTYPE_MAX, (VertexBufferHandle)f, &out->max) != AC_SUCCESS ||
        acDeviceReduceScalDouble(d, s, RTYPE_RMS, (VertexBufferHandle)f, &sum) != AC_SUCCESS)
        return false;
    out->rms = std::sqrt(sum / ncells);
    return true;
}

bool
velocity_stats(Device d, Stream s, double ncells, Stats* out)
{
    double sum = 0;
    if (acDeviceReduceVecDouble(d, s, RTYPE_MIN, VTXBUF_UUX, VTXBUF_UUY, VTXBUF_UUZ, &out->min) !=
            AC_SUCCESS ||
        acDeviceReduceVecDouble(d, s, RTYPE_MAX, VTXBUF_UUX, VTXBUF_UUY, VTXBUF_UUZ, &out->max) !=
            AC_SUCCESS ||
        acDeviceReduceVecDouble(d, s, RTYPE_RMS, VTXBUF_UUX, VTXBUF_UUY, VTXBUF_UUZ, &sum) !=
            AC_SUCCESS)
        return false;
    out->rms = std::sqrt(sum / ncells);
    return true;
}

} // namespace

extern "C" {

AcResult
acHostWriteMeshInfo(const AcMeshInfo info, const char* dir)
{
    FILE* f = fopen(in_dir(dir, "mesh_info.list").c_str(), "w");
    if (!f)
        return AC_FAILURE;
    const unsigned int one = 1;
    const int endian       = (int)*(const unsigned char*)&one; // 1: little endian, 0: big
    fprintf(f, "size_t %s %lu \n", "AcRealSize", (unsigned long)sizeof(AcReal));
    fprintf(f, "int %s %i \n", "endian", endian);
    for (int i = 0; i < NUM_INT_PARAMS; ++i)
        fprintf(f, "int %s %d\n", intparam_names[i], info.int_params[i]);
    for (int i = 0; i < NUM_INT3_PARAMS; ++i)
        fprintf(f, "int3 %s  %d %d %d\n", int3param_names[i], info.int3_params[i].x,
                info.int3_params[i].y, info.int3_params[i].z);
    for (int i = 0; i < NUM_REAL_PARAMS; ++i)
        fprintf(f, "real %s %g\n", realparam_names[i], (double)info.real_params[i]);
    for (int i = 0; i < NUM_REAL3_PARAMS; ++i)
        fprintf(f, "real3 %s  %g %g %g\n", real3param_names[i], (double)info.real3_params[i].x,
                (double)info.real3_params[i].y, (double)info.real3_params[i].z);
    return fclose(f) == 0 ? AC_SUCCESS : AC_FAILURE;
}

AcResult
acHostSaveMeshSnapshot(const AcMesh mesh, const int step, const AcReal t, const char* dir)
{
    const size_t n = acVertexBufferSize(mesh.info);
    for (int w = 0; w < NUM_VTXBUF_HANDLES; ++w) {
        if (!mesh.vertex_buffer[w])
            return AC_FAILURE;
        FILE* f = fopen(mesh_file(dir, w, step).c_str(), "wb");
        if (!f)
            return AC_FAILURE;
        const bool ok = fwrite(&t, sizeof(AcReal), 1, f) == 1 &&
                        fwrite(mesh.vertex_buffer[w], sizeof(AcReal), n, f) == n;
        if (fclose(f) != 0 || !ok)
            return AC_FAILURE;
    }
    return AC_SUCCESS;
}

AcResult
acHostReadMeshSnapshot(AcMesh mesh, const int step, AcReal* t, const char* dir)
{
    const size_t n = acVertexBufferSize(mesh.info);
    for (int w = 0; w < NUM_VTXBUF_HANDLES; ++w) {
        if (!mesh.vertex_buffer[w])
            return AC_FAILURE;
        FILE* f = fopen(mesh_file(dir, w, step).c_str(), "rb");
        if (!f)
            return AC_FAILURE;
        AcReal stamp = 0;
        bool ok      = fread(&stamp, sizeof(AcReal), 1, f) == 1 &&
                  fread(mesh.vertex_buffer[w], sizeof(AcReal), n, f) == n;
        ok = ok && fgetc(f) == EOF; // exactly (1 + m^3) values of this library's AcReal
        fclose(f);
        if (!ok)
            return AC_FAILURE;
        if (t)
            *t = stamp;
    }
    return AC_SUCCESS;
}

AcResult
acTimeseriesBegin(const char* path, const int start_step)
{
    FILE* f = fopen(path, "a");
    if (!f)
        return AC_FAILURE;
    if (start_step == 0) {
        fprintf(f, "step  t_step  dt  uu_total_min  uu_total_rms  uu_total_max  ");
        for (int i = 0; i < NUM_VTXBUF_HANDLES; ++i)
            fprintf(f, "%s_min  %s_rms  %s_max  ", vtxbuf_names[i], vtxbuf_names[i],
                    vtxbuf_names[i]);
    }
    fprintf(f, "\n");
    return fclose(f) == 0 ? AC_SUCCESS : AC_FAILURE;
}

AcResult
acDeviceTimeseriesAppend(const Device device, const Stream stream, const int step,
                         const AcReal dt, const AcReal t, const char* path, int* found_nan)
{
    AcMeshInfo info;
    if (acDeviceGetMeshInfo(device, &info) != AC_SUCCESS)
        return AC_FAILURE;
    const double ncells = (double)info.int_params[AC_nx] * info.int_params[AC_ny] *
                          info.int_params[AC_nz];
    Stats uu, fs[NUM_VTXBUF_HANDLES];
    if (!velocity_stats(device, stream, ncells, &uu))
        return AC_FAILURE;
    for (int i = 0; i < NUM_VTXBUF_HANDLES; ++i)
        if (!field_stats(device, stream, ncells, i, &fs[i]))
            return AC_FAILURE;
    FILE* f = fopen(path, "a");
    if (!f)
        return AC_FAILURE;
    fprintf(f, "%d %e %e %e %e %e ", step, (double)t, (double)dt, uu.min, uu.rms, uu.max);
    int nan = 0;
    for (int i = 0; i < NUM_VTXBUF_HANDLES; ++i) {
        fprintf(f, "%e %e %e ", fs[i].min, fs[i].rms, fs[i].max);
        nan |= std::isnan(fs[i].min) || std::isnan(fs[i].rms) || std::isnan(fs[i].max);
    }
    fprintf(f, "\n");
    if (found_nan && nan)
        *found_nan = 1;
    return fclose(f) == 0 ? AC_SUCCESS : AC_FAILURE;
}

AcResult
acDeviceSaveMeshSnapshot(const Device device, const Stream stream, const int step,
                         const AcReal t, const char* dir)
{
    AcMeshInfo info;
    if (acDeviceGetMeshInfo(device, &info) != AC_SUCCESS)
        return AC_FAILURE;
    const int3 m = mk3(info.int_params[AC_mx], info.int_params[AC_my], info.int_params[AC_mz]);
    AcMesh host;
    if (acHostMeshCreate(info, &host) != AC_SUCCESS)
        return AC_FAILURE;
    AcResult r = acDeviceGeneralBoundconds(device, stream, mk3(0, 0, 0), m, info, mk3(-1, -1, -1));
    if (r == AC_SUCCESS)
        r = acDeviceStoreMesh(device, stream, &host);
    if (r == AC_SUCCESS)
        r = acDeviceSynchronizeStream(device, stream);
    if (r == AC_SUCCESS)
        r = acHostSaveMeshSnapshot(host, step, t, dir);
    acHostMeshDestroy(&host);
    return r;
}

AcResult
acDeviceLoadMeshSnapshot(const Device device, const Stream stream, const int step, AcReal* t,
                         const char* dir)
{
    AcMeshInfo info;
    if (acDeviceGetMeshInfo(device, &info) != AC_SUCCESS)
        return AC_FAILURE;
    AcMesh host;
    if (acHostMeshCreate(info, &host) != AC_SUCCESS)
        return AC_FAILURE;
    AcResult r = acHostReadMeshSnapshot(host, step, t, dir);
    if (r == AC_SUCCESS)
        r = acDeviceLoadMesh(device, stream, host);
    if (r == AC_SUCCESS)
        r = acDeviceSynchronizeStream(device, stream);
    acHostMeshDestroy(&host);
    return r;
}

AcResult
acDeviceRunSimulation(const Device device, const char* dir, AcSimulationResult* result)
{
    AcMeshInfo info;
    if (acDeviceGetMeshInfo(device, &info) != AC_SUCCESS)
        return AC_FAILURE;
    const Stream s          = STREAM_DEFAULT;
    const int start_step    = info.int_params[AC_start_step];
    const int max_steps     = info.int_params[AC_max_steps];
    const int save_steps    = info.int_params[AC_save_steps];
    const int bin_steps     = info.int_params[AC_bin_steps];
    const AcReal max_time   = info.real_params[AC_max_time];
    const AcReal bin_save_t = info.real_params[AC_bin_save_t];
    if (save_steps <= 0 || bin_steps <= 0)
        return AC_FAILURE;
    const std::string ts = in_dir(dir, "timeseries.ts");

    // resume from the snapshot of step start_step, or start from the loaded state
    AcReal t = 0;
    if (start_step > 0 && acDeviceLoadMeshSnapshot(device, s, start_step, &t, dir) != AC_SUCCESS)
        return AC_FAILURE;
    if (acTimeseriesBegin(ts.c_str(), start_step) != AC_SUCCESS ||
        acHostWriteMeshInfo(info, dir) != AC_SUCCESS)
        return AC_FAILURE;
    int found_nan = 0;
    if (start_step == 0) {
        if (acDeviceTimeseriesAppend(device, s, 0, 0, t, ts.c_str(), &found_nan) != AC_SUCCESS ||
            acDeviceSaveMeshSnapshot(device, s, 0, t, dir) != AC_SUCCESS)
            return AC_FAILURE;
    }

    AcReal bin_crit_t = bin_save_t;
    AcReal dt_typical = 0;
    int dtcounter = 0, last = start_step, stop = AC_STOP_MAX_STEPS;
    for (int i = start_step + 1; i < max_steps; ++i) {
        double dtd = 0;
        if (acDeviceTimestepCFL(device, s, &dtd) != AC_SUCCESS)
            return AC_FAILURE;
        const AcReal dt = (AcReal)dtd;
        if (acDeviceIntegrateStepGBC(device, s, info, dt) != AC_SUCCESS)
            return AC_FAILURE;
        t += dt;
        last = i;
        if (i < start_step + 100)
            dt_typical = dt;
        if (i % save_steps == 0 &&
            acDeviceTimeseriesAppend(device, s, i, dt, t, ts.c_str(), &found_nan) != AC_SUCCESS)
            return AC_FAILURE;
        if (i % bin_steps == 0 || t >= bin_crit_t) {
            if (acDeviceSaveMeshSnapshot(device, s, i, t, dir) != AC_SUCCESS)
                return AC_FAILURE;
            bin_crit_t += bin_save_t;
        }
        bool end = false;
        if (max_time > AcReal(0) && t >= max_time) {
            stop = AC_STOP_MAX_TIME;
            break; // the reference stops here without another snapshot
        }
        if (dt < dt_typical / AcReal(1e5)) {
            if (dtcounter > 10) {
                stop = AC_STOP_SMALL_DT;
                end  = true;
            }
            else
                ++dtcounter;
        }
        else
            dtcounter = 0;
        if (!end && found_nan) {
            stop = AC_STOP_NAN;
            end  = true;
        }
        if (!end && access(in_dir(dir, "STOP").c_str(), F_OK) == 0) {
            stop = AC_STOP_FILE;
            end  = true;
        }
        if (end) {
            if (acDeviceSaveMeshSnapshot(device, s, i, t, dir) != AC_SUCCESS)
                return AC_FAILURE;
            break;
        }
    }
    if (result) {
        result->last_step = last;
        result->t         = t;
        result->stop      = stop;
    }
    return AC_SUCCESS;
}

} // extern "C"
