/*
 * fvc_io.c - ompVC-compatible parameter decks and binary formats.
 *
 * Deck: positional "NAME = VALUE" entries, '#' comment lines, names ignored,
 * 97 entries in the order of ompVC/parameters.c:30-138 (reader semantics of
 * parameters.c:258-292).  Decks written before Num_threads existed carry 96
 * entries; they are detected by the name of the 39th entry and read with
 * Num_threads = 0 (the reference would mis-parse them).
 * Binary formats: ompVC/mesh.c:77-94 (_Cartesian.bin), mesh.c:141-205
 * (_coordinates / _connectivity / _adjacency), velocity.c:126-141,184-211
 * (_vel.<n>.bin).
 */
#include "fvc_host.h"

#include <ctype.h>
#include <stddef.h>
#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

static void seterr(char *err, int errlen, const char *fmt, ...) {
    if (!err || errlen <= 0) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, (size_t)errlen, fmt, ap);
    va_end(ap);
}

/* ------------------------------------------------------------------------ */
/* Deck                                                                      */
/* ------------------------------------------------------------------------ */
typedef enum { KS, KI, KD } ktype;
typedef struct {
    const char *name;
    ktype type;
    size_t off;
} kdesc;

#define KE(n, t, f) {n, t, offsetof(fvc_deck, f)}
static const kdesc deck_keys[97] = {
    KE("Path_Data", KS, Path_Data), KE("Path_Output", KS, Path_Output),
    KE("Dimensions", KI, Dimensions), KE("Data_MeshType", KI, Data_MeshType),
    KE("Data_InFilePrefix", KS, Data_InFilePrefix), KE("Data_SuffixTMin", KI, Data_SuffixTMin),
    KE("Data_SuffixTDelta", KI, Data_SuffixTDelta), KE("Data_TRes", KI, Data_TRes),
    KE("Data_TDelta", KD, Data_TDelta), KE("Data_TMin", KD, Data_TMin),
    KE("Data_TPeriodic", KI, Data_TPeriodic),
    KE("Data_MeshBounds.XMin", KD, Bounds[0]), KE("Data_MeshBounds.XMax", KD, Bounds[1]),
    KE("Data_MeshBounds.YMin", KD, Bounds[2]), KE("Data_MeshBounds.YMax", KD, Bounds[3]),
    KE("Data_MeshBounds.ZMin", KD, Bounds[4]), KE("Data_MeshBounds.ZMax", KD, Bounds[5]),
    KE("Fluid_Density", KD, Fluid_Density), KE("Fluid_Viscosity", KD, Fluid_Viscosity),
    KE("Output_TStart", KD, Output_TStart), KE("Output_TRes", KI, Output_TRes),
    KE("Output_TDelta", KD, Output_TDelta),
    KE("Int_Type", KI, Int_Type), KE("Int_TimeStep", KD, Int_TimeStep),
    KE("Int_Accuracy", KD, Int_Accuracy), KE("Int_MinTimeStep", KD, Int_MinTimeStep),
    KE("Int_MaxTimeStep", KD, Int_MaxTimeStep), KE("Int_TimeDirection", KI, Int_TimeDirection),
    KE("Int_NormalFlow", KI, Int_NormalFlow), KE("Int_NormalFlowScaling", KD, Int_NormalFlowScaling),
    KE("Int_Extrapolate", KI, Int_Extrapolate),
    KE("Particle_Radius", KD, Particle_Radius), KE("Particle_Density", KD, Particle_Density),
    KE("Particle_ICType", KI, Particle_ICType),
    KE("Gravity_Vector[0]", KD, Gravity[0]), KE("Gravity_Vector[1]", KD, Gravity[1]),
    KE("Gravity_Vector[2]", KD, Gravity[2]),
    KE("LocalSearchChecking", KI, LocalSearchChecking), KE("Num_threads", KI, Num_threads),
    KE("FTLE_Compute", KI, FTLE_Compute), KE("FTLE_GenerateMesh", KI, FTLE_GenerateMesh),
    KE("FTLE_ICFile", KS, FTLE_ICFile),
    KE("FTLE_MeshBounds.XMin", KD, FTLE_Bounds[0]), KE("FTLE_MeshBounds.XMax", KD, FTLE_Bounds[1]),
    KE("FTLE_MeshBounds.YMin", KD, FTLE_Bounds[2]), KE("FTLE_MeshBounds.YMax", KD, FTLE_Bounds[3]),
    KE("FTLE_MeshBounds.ZMin", KD, FTLE_Bounds[4]), KE("FTLE_MeshBounds.ZMax", KD, FTLE_Bounds[5]),
    KE("FTLE_MeshBounds.XRes", KI, FTLE_Res[0]), KE("FTLE_MeshBounds.YRes", KI, FTLE_Res[1]),
    KE("FTLE_MeshBounds.ZRes", KI, FTLE_Res[2]),
    KE("FTLE_IntTLength", KD, FTLE_IntTLength), KE("FTLE_ComputeVariation", KI, FTLE_ComputeVariation),
    KE("FTLE_VariationOutFreq", KI, FTLE_VariationOutFreq), KE("FTLE_OutFilePrefix", KS, FTLE_OutFilePrefix),
    KE("Trace_Compute", KI, Trace_Compute), KE("Trace_ReleaseStrategy", KI, Trace_ReleaseStrategy),
    KE("Trace_ReleaseTMax", KD, Trace_ReleaseTMax), KE("Trace_GenerateMesh", KI, Trace_GenerateMesh),
    KE("Trace_InFile", KS, Trace_InFile), KE("Trace_MultipleInFiles", KI, Trace_MultipleInFiles),
    KE("Trace_InFileFormat", KI, Trace_InFileFormat), KE("Trace_OutFilePrefix", KS, Trace_OutFilePrefix),
    KE("Trace_NumLaunchTimes", KI, Trace_NumLaunchTimes),
    KE("Trace_LaunchTimeSpacing", KD, Trace_LaunchTimeSpacing),
    KE("Trace_IntTLength", KD, Trace_IntTLength), KE("Trace_AlwaysOutput", KI, Trace_AlwaysOutput),
    KE("Trace_MeshBounds.XMin", KD, Trace_Bounds[0]), KE("Trace_MeshBounds.XMax", KD, Trace_Bounds[1]),
    KE("Trace_MeshBounds.YMin", KD, Trace_Bounds[2]), KE("Trace_MeshBounds.YMax", KD, Trace_Bounds[3]),
    KE("Trace_MeshBounds.ZMin", KD, Trace_Bounds[4]), KE("Trace_MeshBounds.ZMax", KD, Trace_Bounds[5]),
    KE("Trace_MeshBounds.XRes", KI, Trace_Res[0]), KE("Trace_MeshBounds.YRes", KI, Trace_Res[1]),
    KE("Trace_MeshBounds.ZRes", KI, Trace_Res[2]),
    KE("Trace_APCompute", KI, Trace_APCompute), KE("Trace_CETCompute", KI, Trace_CETCompute),
    KE("Trace_CETAuxillaryMesh", KI, Trace_CETAuxillaryMesh), KE("Trace_CETMeshPrefix", KS, Trace_CETMeshPrefix),
    KE("Trace_CETSubsteps", KI, Trace_CETSubsteps), KE("Trace_RTCompute", KI, Trace_RTCompute),
    KE("Trace_RTOutFilePrefix", KS, Trace_RTOutFilePrefix),
    KE("VelOut_Compute", KI, VelOut_Compute), KE("VelOut_GenerateMesh", KI, VelOut_GenerateMesh),
    KE("VelOut_InFile", KS, VelOut_InFile), KE("VelOut_InFileFormat", KI, VelOut_InFileFormat),
    KE("VelOut_FilePrefix", KS, VelOut_FilePrefix),
    KE("VelOut_MeshBounds.XMin", KD, VelOut_Bounds[0]), KE("VelOut_MeshBounds.XMax", KD, VelOut_Bounds[1]),
    KE("VelOut_MeshBounds.YMin", KD, VelOut_Bounds[2]), KE("VelOut_MeshBounds.YMax", KD, VelOut_Bounds[3]),
    KE("VelOut_MeshBounds.ZMin", KD, VelOut_Bounds[4]), KE("VelOut_MeshBounds.ZMax", KD, VelOut_Bounds[5]),
    KE("VelOut_MeshBounds.XRes", KI, VelOut_Res[0]), KE("VelOut_MeshBounds.YRes", KI, VelOut_Res[1]),
    KE("VelOut_MeshBounds.ZRes", KI, VelOut_Res[2]),
};
#define NUM_THREADS_INDEX 38

void fvc_deck_defaults(fvc_deck *d) {
    memset(d, 0, sizeof *d);
    strcpy(d->Path_Data, "pwd");
    strcpy(d->Path_Output, "pwd");
    d->Dimensions = 3;
    d->Data_MeshType = 1;
    strcpy(d->Data_InFilePrefix, "data");
    d->Data_SuffixTDelta = 1;
    d->Data_TRes = 2;
    d->Data_TDelta = 1.0;
    d->Fluid_Density = 1.0;
    d->Fluid_Viscosity = 1.0;
    d->Output_TRes = 1;
    d->Output_TDelta = 1.0;
    d->Int_Type = 1;
    d->Int_TimeStep = 1e-3;
    d->Int_Accuracy = 1e-4;
    d->Int_MinTimeStep = 1e-4;
    d->Int_MaxTimeStep = 1e-3;
    d->Int_TimeDirection = 1;
    d->Int_NormalFlowScaling = 1.0;
    d->Particle_Density = 1.0;
    strcpy(d->FTLE_ICFile, "ftle_ic.bin");
    d->FTLE_Res[0] = d->FTLE_Res[1] = d->FTLE_Res[2] = 2;
    d->FTLE_IntTLength = 1.0;
    d->FTLE_VariationOutFreq = 1;
    strcpy(d->FTLE_OutFilePrefix, "ftle");
    strcpy(d->Trace_InFile, "seeds.bin");
    d->Trace_InFileFormat = 4;
    strcpy(d->Trace_OutFilePrefix, "tracer");
    d->Trace_NumLaunchTimes = 1;
    d->Trace_LaunchTimeSpacing = 1.0;
    d->Trace_IntTLength = 1.0;
    d->Trace_Res[0] = d->Trace_Res[1] = d->Trace_Res[2] = 1;
    strcpy(d->Trace_CETMeshPrefix, "aux");
    d->Trace_CETSubsteps = 2;
    strcpy(d->Trace_RTOutFilePrefix, "rt");
    strcpy(d->VelOut_InFile, "velin.vtk");
    d->VelOut_InFileFormat = 1;
    strcpy(d->VelOut_FilePrefix, "velout");
    d->VelOut_Res[0] = d->VelOut_Res[1] = d->VelOut_Res[2] = 1;
}

/* token reader with '#'-to-end-of-line comments between entries */
typedef struct {
    FILE *f;
} tokrd;

static int next_token(tokrd *r, char *buf, int buflen, int allow_comment) {
    int c;
    for (;;) {
        c = fgetc(r->f);
        if (c == EOF) return 0;
        if (isspace(c)) continue;
        if (c == '#' && allow_comment) {
            while (c != '\n' && c != EOF) c = fgetc(r->f);
            continue;
        }
        break;
    }
    int n = 0;
    while (c != EOF && !isspace(c)) {
        if (n < buflen - 1) buf[n++] = (char)c;
        c = fgetc(r->f);
    }
    buf[n] = 0;
    return 1;
}

int fvc_deck_read(const char *path, fvc_deck *d, char *err, int errlen) {
    fvc_deck_defaults(d);
    FILE *f = fopen(path, "r");
    if (!f) {
        seterr(err, errlen, "Could not open input file %s", path);
        return -1;
    }
    tokrd r = {f};
    char name[FVC_STR * 2], eq[FVC_STR * 2], val[FVC_STR * 2];
    int k = 0;
    while (k < 97) {
        if (!next_token(&r, name, sizeof name, 1)) break;
        if (k == NUM_THREADS_INDEX && strcmp(name, "FTLE_Compute") == 0) {
            d->legacy_96 = 1; /* 96-key deck: Num_threads absent */
            d->Num_threads = 0;
            k++;
        }
        if (!next_token(&r, eq, sizeof eq, 0) || !next_token(&r, val, sizeof val, 0)) break;
        const kdesc *kd = &deck_keys[k];
        char *p = (char *)d + kd->off;
        char *end = NULL;
        /* the reference reads with fscanf into fixed buffers and leaves an
         * unparsable number unset; both are refused here */
        if (kd->type == KS) {
            if (strlen(val) >= FVC_STR) {
                fclose(f);
                seterr(err, errlen, "Value of %s is longer than %d characters", kd->name, FVC_STR - 1);
                return -3;
            }
            strcpy(p, val);
        } else if (kd->type == KI) {
            long v = strtol(val, &end, 10);
            if (end == val || *end || v < INT32_MIN || v > INT32_MAX) {
                fclose(f);
                seterr(err, errlen, "Could not read an integer for %s from \"%s\"", kd->name, val);
                return -3;
            }
            *(int32_t *)p = (int32_t)v;
        } else {
            double v = strtod(val, &end);
            if (end == val || *end) {
                fclose(f);
                seterr(err, errlen, "Could not read a number for %s from \"%s\"", kd->name, val);
                return -3;
            }
            *(double *)p = v;
        }
        k++;
    }
    fclose(f);
    if (k < 97) {
        seterr(err, errlen, "deck %s ended after %d of 97 entries", path, k);
        return -2;
    }
    return 0;
}

int fvc_deck_write(const char *path, const fvc_deck *d) {
    FILE *f = fopen(path, "w");
    if (!f) return -1;
    fprintf(f, "# flowVC parameter deck (97 positional entries)\n");
    for (int k = 0; k < 97; k++) {
        const kdesc *kd = &deck_keys[k];
        const char *p = (const char *)d + kd->off;
        if (kd->type == KS)
            fprintf(f, "%s = %s\n", kd->name, p);
        else if (kd->type == KI)
            fprintf(f, "%s = %d\n", kd->name, *(const int32_t *)p);
        else
            fprintf(f, "%s = %.17g\n", kd->name, *(const double *)p);
    }
    fclose(f);
    return 0;
}

int fvc_deck_check(const fvc_deck *d, char *err, int errlen) {
#define FAIL(...) do { seterr(err, errlen, __VA_ARGS__); return -1; } while (0)
    if (d->Dimensions != 2 && d->Dimensions != 3) FAIL("Dimensions must be equal to 2 or 3");
    if (d->Data_MeshType != 0 && d->Data_MeshType != 1) FAIL("Unsupported Data_MeshType");
    if (d->Data_TRes < 2) FAIL("Data_TRes must be 2 or greater.");
    if (d->Data_TDelta <= 0) FAIL("Data_TDelta <= 0");
    if (d->Data_TPeriodic != 0 && d->Data_TPeriodic != 1) FAIL("Data_TPeriodic should be 0 or 1");
    if (d->Bounds[0] >= d->Bounds[1]) FAIL("Data_MeshBounds.XMin >= Data_MeshBounds.XMax");
    if (d->Bounds[2] >= d->Bounds[3]) FAIL("Data_MeshBounds.YMin >= Data_MeshBounds.YMax");
    if (d->Bounds[4] > d->Bounds[5]) FAIL("Data_MeshBounds.ZMin > Data_MeshBounds.ZMax");
    if (d->Output_TRes < 1) FAIL("Output_TRes < 1");
    if (d->Output_TDelta < FVC_TINY && d->Output_TRes > 1) FAIL("Output_TDelta must be positive");
    if (d->Int_TimeStep < FVC_TINY && d->Int_Type != 2) FAIL("Int_TimeStep must be positive");
    if ((d->Int_MinTimeStep < FVC_TINY || d->Int_MaxTimeStep < FVC_TINY || d->Int_Accuracy < FVC_TINY) &&
        d->Int_Type == 2) /* parameters.c:462-463 */
        FAIL("RKF integration parameters must be positive");
    if (d->Int_TimeDirection != 1 && d->Int_TimeDirection != -1)
        FAIL("Unsupported Int_TimeDirection value, must be -1 or 1");
    if (d->FTLE_Compute && d->Trace_Compute)
        FAIL("Both FTLE_Compute = 1 and Trace_Compute = 1 cannot be handled concurrently");
    if (d->FTLE_Compute) {
        if (d->FTLE_Bounds[0] >= d->FTLE_Bounds[1]) FAIL("FTLE_CartMesh.XMin >= FTLE_CartMesh.XMax");
        if (d->FTLE_Bounds[2] >= d->FTLE_Bounds[3]) FAIL("FTLE_CartMesh.YMin >= FTLE_CartMesh.YMax");
        if (d->FTLE_Bounds[4] > d->FTLE_Bounds[5]) FAIL("FTLE_CartMesh.ZMin > FTLE_CartMesh.ZMax");
        if (d->FTLE_Res[0] < 2) FAIL("FTLE_CartMesh.XRes < 2");
        if (d->FTLE_Res[1] < 2) FAIL("FTLE_CartMesh.YRes < 2");
        if (d->Dimensions == 3 && d->FTLE_Res[2] < 2) FAIL("FTLE_CartMesh.ZRes insufficient");
        if (d->FTLE_IntTLength < FVC_TINY) FAIL("FTLE_IntTLength must be positive");
    }
    if (d->VelOut_Compute) { /* parameters.c:546-563 */
        if (d->VelOut_GenerateMesh) {
            for (int a = 0; a < 3; a++) {
                if (d->VelOut_Bounds[2 * a] > d->VelOut_Bounds[2 * a + 1]) FAIL("VelOut_CartMesh Min > Max");
                if (d->VelOut_Res[a] < 1) FAIL("VelOut_CartMesh Res < 1");
            }
        }
        if (d->VelOut_InFileFormat != 1 && d->VelOut_InFileFormat != 2 && d->VelOut_InFileFormat != 3)
            FAIL("Unsupported value for VelOut_InFileFormat");
    }
    if (d->Trace_Compute) {
        if (d->Trace_ReleaseStrategy == 1) { /* parameters.c:500-511 */
            if (d->Trace_AlwaysOutput)
                fprintf(stderr, "\nWarning: Trace_AlwaysOutput not supported for Trace_ReleaseStrategy = 1\n\n");
            if (d->Int_TimeDirection < 0) FAIL("Staggered tracer release not supported for backward time integration");
            if (d->Trace_ReleaseTMax < 0) FAIL("Trace_ReleaseTMax < 0");
            if (d->Dimensions == 2) FAIL("Staggered release not supported for 2D analysis");
            if (d->Trace_InFileFormat != 4) FAIL("Trace_InFileFormat must be 4 for Trace_ReleaseStrategy = 1");
        }
        if (d->Trace_GenerateMesh) {
            for (int a = 0; a < 3; a++) {
                if (d->Trace_Bounds[2 * a] > d->Trace_Bounds[2 * a + 1]) FAIL("Trace_CartMesh Min > Max");
                if (d->Trace_Res[a] < 1) FAIL("Trace_CartMesh Res < 1");
            }
        } else if (d->Trace_InFileFormat != 0 && d->Trace_InFileFormat != 1 && d->Trace_InFileFormat != 4 &&
                   d->Trace_InFileFormat != 2 && d->Trace_InFileFormat != 3) {
            FAIL("Unrecognized value for Trace_InFileFormat");
        }
        if (d->Trace_NumLaunchTimes < 1) FAIL("Trace_NumLaunchTimes < 1");
        if (d->Trace_LaunchTimeSpacing < FVC_TINY && d->Trace_NumLaunchTimes > 1)
            FAIL("Trace_LaunchTimeSpacing must be positive");
        if (d->Trace_IntTLength < 0) FAIL("Trace_IntTLength < 0");
        if (d->Trace_InFileFormat == 0 && d->Trace_RTCompute && !d->Trace_GenerateMesh) /* parameters.c:543-544 */
            fprintf(stderr, "  Warning: Trace_InFileFormat = 0 not recommended when Trace_RTCompute = 1 as connectivity "
                            "data will be lost.\n");
        if (d->Trace_CETCompute) { /* parameters.c:535-542 */
            if (d->Trace_RTCompute) FAIL("Cannot compute exposure times and residence times concurrently");
            if (d->Trace_CETSubsteps < 2) FAIL("Trace_CETSubsteps must be greater than 1");
        }
        if ((d->Trace_RTCompute || d->Trace_CETCompute) && d->Int_TimeDirection < 0)
            FAIL("Residence/Exposure time computations not supported for backward time integrations");
    }
    return 0;
#undef FAIL
}

static void fix_path(char *p) {
    if (strcmp(p, "pwd") == 0) {
        strcpy(p, "./");
        return;
    }
    size_t n = strlen(p);
    if (n > 0 && p[n - 1] != '/' && n + 1 < FVC_STR) {
        p[n] = '/';
        p[n + 1] = 0;
    }
}

int fvc_deck_derive(fvc_deck *d, char *err, int errlen) {
    int dir = d->Int_TimeDirection;
    d->Data_TMax = d->Data_TMin + (d->Data_TRes - 1) * d->Data_TDelta;
    if ((d->Output_TStart < d->Data_TMin || d->Output_TStart > d->Data_TMax) && !d->Data_TPeriodic) {
        seterr(err, errlen, "Output_TStart outside data range");
        return -1;
    }
    d->Output_TEnd = d->Output_TStart + (d->Output_TRes - 1) * dir * d->Output_TDelta;
    if ((d->Output_TEnd > d->Data_TMax || d->Output_TEnd < d->Data_TMin) && !d->Data_TPeriodic) {
        seterr(err, errlen, "Output_TEnd outside data range");
        return -1;
    }
    if (d->FTLE_Compute) {
        fvc_cart *c = &d->FTLE_Cart;
        memset(c, 0, sizeof *c);
        c->XMin = d->FTLE_Bounds[0]; c->XMax = d->FTLE_Bounds[1]; c->XRes = d->FTLE_Res[0];
        c->YMin = d->FTLE_Bounds[2]; c->YMax = d->FTLE_Bounds[3]; c->YRes = d->FTLE_Res[1];
        c->ZMin = d->FTLE_Bounds[4]; c->ZMax = d->FTLE_Bounds[5]; c->ZRes = d->FTLE_Res[2];
        c->XDelta = (c->XMax - c->XMin) / (c->XRes - 1);
        c->YDelta = (c->YMax - c->YMin) / (c->YRes - 1);
        if (d->Dimensions == 3)
            c->ZDelta = (c->ZMax - c->ZMin) / (c->ZRes - 1);
        else {
            c->ZMin = 0; c->ZMax = 0; c->ZDelta = 0; c->ZRes = 1;
        }
        double a = d->Output_TStart, b = d->Output_TEnd + dir * d->FTLE_IntTLength;
        double tminreq = a < b ? a : b, tmaxreq = a > b ? a : b;
        if (tminreq < d->Data_TMin && !d->Data_TPeriodic) tminreq = d->Data_TMin;
        if (tmaxreq > d->Data_TMax && !d->Data_TPeriodic) tmaxreq = d->Data_TMax;
        d->Data_FirstFrame = (int32_t)(dir > 0 ? floor((tminreq - d->Data_TMin) / d->Data_TDelta)
                                               : ceil((tmaxreq - d->Data_TMin) / d->Data_TDelta));
        d->Data_LastFrame = (int32_t)(dir > 0 ? ceil((tmaxreq - d->Data_TMin) / d->Data_TDelta)
                                              : floor((tminreq - d->Data_TMin) / d->Data_TDelta));
    } else {
        d->Data_FirstFrame = (int32_t)(dir > 0 ? floor((d->Output_TStart - d->Data_TMin) / d->Data_TDelta)
                                               : ceil((d->Output_TStart - d->Data_TMin) / d->Data_TDelta));
        d->Data_LastFrame = (int32_t)(dir > 0 ? ceil((d->Output_TEnd - d->Data_TMin) / d->Data_TDelta)
                                              : floor((d->Output_TEnd - d->Data_TMin) / d->Data_TDelta));
    }
    if (d->Trace_GenerateMesh) {
        fvc_cart *c = &d->Trace_Cart;
        memset(c, 0, sizeof *c);
        c->XMin = d->Trace_Bounds[0]; c->XMax = d->Trace_Bounds[1]; c->XRes = d->Trace_Res[0];
        c->YMin = d->Trace_Bounds[2]; c->YMax = d->Trace_Bounds[3]; c->YRes = d->Trace_Res[1];
        c->ZMin = d->Trace_Bounds[4]; c->ZMax = d->Trace_Bounds[5]; c->ZRes = d->Trace_Res[2];
        c->XDelta = c->XRes == 1 ? 0.0 : (c->XMax - c->XMin) / ((double)c->XRes - 1.0);
        c->YDelta = c->YRes == 1 ? 0.0 : (c->YMax - c->YMin) / ((double)c->YRes - 1.0);
        if (c->ZRes == 1)
            c->ZDelta = 0.0;
        else if (d->Dimensions == 3)
            c->ZDelta = (c->ZMax - c->ZMin) / ((double)c->ZRes - 1.0);
        else {
            c->ZMin = 0; c->ZMax = 0; c->ZDelta = 0; c->ZRes = 1;
        }
    }
    fix_path(d->Path_Output);
    fix_path(d->Path_Data);
    if (d->Particle_Radius < 0) {
        seterr(err, errlen, "Particle_Radius cannot be negative");
        return -1;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Binary files                                                              */
/* ------------------------------------------------------------------------ */
int fvc_write_cart_header(const char *path, const double mn[3], const double mx[3], const int32_t res[3]) {
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    int ok = 1;
    for (int a = 0; a < 3; a++) {
        ok &= fwrite(&mn[a], 8, 1, f) == 1;
        ok &= fwrite(&mx[a], 8, 1, f) == 1;
        ok &= fwrite(&res[a], 4, 1, f) == 1;
    }
    fclose(f);
    return ok ? 0 : -1;
}

int fvc_read_cart_header(const char *path, double mn[3], double mx[3], int32_t res[3]) {
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    int ok = 1;
    for (int a = 0; a < 3; a++) {
        ok &= fread(&mn[a], 8, 1, f) == 1;
        ok &= fread(&mx[a], 8, 1, f) == 1;
        ok &= fread(&res[a], 4, 1, f) == 1;
    }
    fclose(f);
    return ok ? 0 : -2;
}

int fvc_write_vel_slice(const char *path, double t, const double *uvw, int64_t n) {
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    int ok = fwrite(&t, 8, 1, f) == 1 && fwrite(uvw, 8, (size_t)(3 * n), f) == (size_t)(3 * n);
    fclose(f);
    return ok ? 0 : -1;
}

int fvc_read_vel_slice(const char *path, double *t, double *uvw, int64_t n) {
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    int ok = fread(t, 8, 1, f) == 1 && fread(uvw, 8, (size_t)(3 * n), f) == (size_t)(3 * n);
    fclose(f);
    return ok ? 0 : -2;
}

int fvc_write_unstruct_mesh(const char *dir, const char *prefix, int64_t nn, const double *xyz,
                            int64_t ne, const int32_t *conn, const int32_t *adj) {
    char path[2 * FVC_STR + 32];
    int32_t n32 = (int32_t)nn, e32 = (int32_t)ne;
    FILE *f;
    snprintf(path, sizeof path, "%s/%s_coordinates.bin", dir, prefix);
    if (!(f = fopen(path, "wb"))) return -1;
    fwrite(&n32, 4, 1, f);
    fwrite(xyz, 8, (size_t)(3 * nn), f);
    fclose(f);
    snprintf(path, sizeof path, "%s/%s_connectivity.bin", dir, prefix);
    if (!(f = fopen(path, "wb"))) return -1;
    fwrite(&e32, 4, 1, f);
    fwrite(conn, 4, (size_t)(4 * ne), f);
    fclose(f);
    snprintf(path, sizeof path, "%s/%s_adjacency.bin", dir, prefix);
    if (!(f = fopen(path, "wb"))) return -1;
    fwrite(&e32, 4, 1, f);
    fwrite(adj, 4, (size_t)(4 * ne), f);
    fclose(f);
    return 0;
}

int fvc_read_unstruct_counts(const char *dir, const char *prefix, int64_t *nn, int64_t *ne) {
    char path[2 * FVC_STR + 32];
    int32_t v;
    FILE *f;
    snprintf(path, sizeof path, "%s%s_coordinates.bin", dir, prefix);
    if (!(f = fopen(path, "rb"))) return -1;
    if (fread(&v, 4, 1, f) != 1) { fclose(f); return -2; }
    *nn = v;
    fclose(f);
    snprintf(path, sizeof path, "%s%s_connectivity.bin", dir, prefix);
    if (!(f = fopen(path, "rb"))) return -1;
    if (fread(&v, 4, 1, f) != 1) { fclose(f); return -2; }
    *ne = v;
    fclose(f);
    return 0;
}

int fvc_read_unstruct_mesh(const char *dir, const char *prefix, double *xyz, int32_t *conn, int32_t *adj) {
    char path[2 * FVC_STR + 32];
    int32_t nn, ne, ne2;
    FILE *f;
    snprintf(path, sizeof path, "%s%s_coordinates.bin", dir, prefix);
    if (!(f = fopen(path, "rb"))) return -1;
    int ok = fread(&nn, 4, 1, f) == 1 && fread(xyz, 8, (size_t)3 * nn, f) == (size_t)3 * nn;
    fclose(f);
    if (!ok) return -2;
    snprintf(path, sizeof path, "%s%s_connectivity.bin", dir, prefix);
    if (!(f = fopen(path, "rb"))) return -1;
    ok = fread(&ne, 4, 1, f) == 1 && fread(conn, 4, (size_t)4 * ne, f) == (size_t)4 * ne;
    fclose(f);
    if (!ok) return -2;
    snprintf(path, sizeof path, "%s%s_adjacency.bin", dir, prefix);
    if (!(f = fopen(path, "rb"))) return -1;
    ok = fread(&ne2, 4, 1, f) == 1;
    if (ok && ne2 != ne) {
        fclose(f);
        return -3; /* mesh.c:199-200 */
    }
    ok = ok && fread(adj, 4, (size_t)4 * ne, f) == (size_t)4 * ne;
    fclose(f);
    return ok ? 0 : -2;
}
