// C ABI of the uniform-layout AST projection (astg_ast_*), the device side of
// the reference's ast::ASTSolver<d> (include/ast_solver.h:20-673) projection
// path; the declarations and their reference counterparts are in astg.h.
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <new>
#include <string>

#include "../../include/astg.h"
#include "ast_uniform.h"

struct astg_ast {
    int device = 0;
    hipStream_t stream = nullptr;
    std::unique_ptr<astg::AstUniform> ast;
    std::string err;
};

namespace {

int fail(astg_ast* a, int code, const std::string& m) {
    if (a) a->err = m;
    return code;
}

template <typename F>
int guarded(astg_ast* a, F&& f) {
    if (!a) return ASTG_ERR_ARG;
    if (!a->ast || !a->ast->ready()) return fail(a, ASTG_ERR_STATE, "context was not created");
    if (hipSetDevice(a->device) != hipSuccess) return fail(a, ASTG_ERR_HIP, "hipSetDevice failed");
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return fail(a, ASTG_ERR_ARG, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(a, ASTG_ERR_HIP, e.what());
    }
}

}  // namespace

extern "C" {

void astg_ast_params_default(astg_pcg_params* p) {
    if (!p) return;
    astg_pcg_params_default(p);
    // ASTSolver::initialize (ast_solver.h:40): SOLVER_ACCURACY 1e-5, 5000
    // iterations, MIC(0) 0.97, min-diagonal ratio 0.95
    p->tol = 1e-5;
    p->max_iterations = 5000;
    p->mic_tau = 0.97;
    p->min_diag_ratio = 0.95;
    p->ordering = ASTG_ORDER_REFERENCE;
}

int astg_ast_create(astg_ast** out, int device, int dim, const int* res, double dx) {
    if (!out || !res) return ASTG_ERR_ARG;
    *out = nullptr;
    int count = 0;
    if (device < 0 || hipGetDeviceCount(&count) != hipSuccess || device >= count) return ASTG_ERR_HIP;
    auto* a = new (std::nothrow) astg_ast();
    if (!a) return ASTG_ERR_ARG;
    a->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&a->stream) != hipSuccess) {
        delete a;
        return ASTG_ERR_HIP;
    }
    a->ast.reset(new astg::AstUniform(a->stream));
    const int rc = a->ast->init(dim, res, dx, a->err);
    *out = a;  // on failure the caller reads the message, then destroys it
    return rc;
}

int astg_ast_destroy(astg_ast* a) {
    if (!a) return ASTG_OK;
    (void)hipSetDevice(a->device);
    a->ast.reset();
    if (a->stream) (void)hipStreamDestroy(a->stream);
    delete a;
    return ASTG_OK;
}

const char* astg_ast_last_error(const astg_ast* a) { return a ? a->err.c_str() : "null context"; }

int astg_ast_set_layout(astg_ast* a, const uint8_t* tilt_open) {
    return guarded(a, [&] {
        if (!tilt_open) return fail(a, ASTG_ERR_ARG, "tilt_open is null");
        return a->ast->setLayout(tilt_open, a->err);
    });
}

int astg_ast_set_velocity(astg_ast* a, const double* mac_x, const double* mac_y, const double* mac_z,
                          const double* tilt) {
    return guarded(a, [&] {
        const double* m[3] = {mac_x, mac_y, mac_z};
        if (!mac_x || !mac_y || !tilt || (a->ast->geom().d == 3 && !mac_z))
            return fail(a, ASTG_ERR_ARG, "a velocity array is null");
        return a->ast->setVelocity(m, tilt, a->err);
    });
}

int astg_ast_get_velocity(astg_ast* a, double* mac_x, double* mac_y, double* mac_z, double* tilt) {
    return guarded(a, [&] {
        double* m[3] = {mac_x, mac_y, mac_z};
        if (!mac_x || !mac_y || !tilt || (a->ast->geom().d == 3 && !mac_z))
            return fail(a, ASTG_ERR_ARG, "a velocity array is null");
        return a->ast->getVelocity(m, tilt, a->err);
    });
}

int astg_ast_project(astg_ast* a, double dt, double rho, const astg_pcg_params* params, astg_pcg_result* result) {
    return guarded(a, [&] {
        astg_pcg_params prm;
        astg_ast_params_default(&prm);
        if (params) prm = *params;
        astg_pcg_result r;
        std::memset(&r, 0, sizeof(r));
        const int rc = a->ast->project(dt, rho, prm, r, a->err);
        if (result) *result = r;
        return rc;
    });
}

int astg_ast_system(astg_ast* a, int n, long long nnz, int* rowstart, int* col, double* val, double* b, double* p) {
    return guarded(a, [&] {
        if (n != a->ast->n() || nnz != a->ast->nnz())
            return fail(a, ASTG_ERR_ARG, "n / nnz do not match the last projection");
        return a->ast->system(rowstart, col, val, b, p, a->err);
    });
}

int astg_ast_tilt_index(astg_ast* a, int* tilt_index) {
    return guarded(a, [&] {
        if (!tilt_index) return fail(a, ASTG_ERR_ARG, "tilt_index is null");
        return a->ast->tiltIndex(tilt_index, a->err);
    });
}

}  // extern "C"
