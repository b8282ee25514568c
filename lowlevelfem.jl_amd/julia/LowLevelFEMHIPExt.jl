module LowLevelFEMHIPExt
# Package extension of LowLevelFEM.jl that routes stiffnessMatrix, massMatrix,
# loadVector and solveDisplacement of a 3D :Solid problem to libllfem.so
# (C ABI: include/llfem.h).  The structs below are field for field the C
# structs of llfem.h; every ccall's argument types follow the C prototype.

using LowLevelFEM, SparseArrays
import LowLevelFEM: gmsh, Problem, SystemMatrix, VectorField, BoundaryCondition

const lib = get(ENV, "LLFEM_LIB", "libllfem.so")

struct OpDesc                      # llfem_op_desc
    op::Int32; q::Int32
    weights::Ptr{Float64}; N_ref::Ptr{Float64}; dN_ref::Ptr{Float64}; D::Ptr{Float64}
    E::Float64; nu::Float64; rho::Float64
    lumped::Int32; precision::Int32
    element_scale::Ptr{Float64}    # [E] per-element integrand factor, or C_NULL
    accumulate::Int32              # 0: replace the mesh's resident values, 1: add to them
end

mutable struct Stats               # llfem_stats
    kernel_ms::Float64; main_kernel_ms::Float64; total_ms::Float64
    nnz::Int64; elements::Int64; alg_bytes::Float64
    iterations::Int64; refinements::Int64; relres::Float64; alg_flops::Float64; setup_ms::Float64
    rollbacks::Int64; restarts::Int64
    Stats() = new(0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0)
end

struct Request                     # llfem_request
    xyz::Ptr{Float64}; N::Int64; elem_type::Int32; npe::Int32
    conn::Ptr{Int64}; E::Int64; stiffness::OpDesc
    face_type::Int32; nf::Int32; face_conn::Ptr{Int64}; F::Int64
    face_q::Int32; face_weights::Ptr{Float64}; face_N_ref::Ptr{Float64}; face_dN_ref::Ptr{Float64}
    traction::NTuple{3,Float64}; fixed::Ptr{Int64}; nfixed::Int64
    method::Int32; reltol::Float64; maxiter::Int64; mesh_key::UInt64; u_out::Ptr{Float64}
end

mutable struct RequestStats        # llfem_request_stats
    latency_ms::Float64; queue_ms::Float64; service_ms::Float64
    iterations::Int64; updates::Int64; relres::Float64
    mesh_cache_hit::Int32; status::Int32
    rollbacks::Int64; restarts::Int64
    RequestStats() = new(0, 0, 0, 0, 0, 0, 0, 0, 0, 0)
end

const OP_STIFFNESS = Int32(0)
const OP_MASS = Int32(1)
const PCG_JACOBI_FP64 = Int32(0)
const PCG_MIXED_REFINE = Int32(1)
const PCG_AMG_RELIABLE = Int32(4)

check(rc) = rc == 0 || error(unsafe_string(ccall((:llfem_last_error, lib), Cstring, ())))

const CTX = Ref{Ptr{Cvoid}}(C_NULL)
function ctx()
    if CTX[] == C_NULL
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:llfem_ctx_create, lib), Cint, (Cint, Csize_t, Ref{Ptr{Cvoid}}), 0, 0, r))
        CTX[] = r[]
    end
    CTX[]
end

# One device mesh per (problem, volume element type): coordinates of all
# nodes, that type's connectivity, its CSC pattern and element->nnz map,
# built once.  elemTags are the type's element tags in gmsh's order (the
# order of the device mesh's elements); pos maps an element tag to its index.
struct TypeMesh
    et::Int32
    nn::Int32
    order::Int
    handle::Ptr{Cvoid}
    elemTags::Vector{Int64}
    pos::Dict{Int64,Int}
end

const MESHES = Dict{Tuple{String,Int32},TypeMesh}()

function type_mesh(problem::Problem, et::Integer)
    get!(MESHES, (problem.name, Int32(et))) do
        gmsh.model.setCurrent(problem.name)
        _, _, order, nn, _, _ = gmsh.model.mesh.getElementProperties(et)
        elemTags, nodeTags = gmsh.model.mesh.getElementsByType(et)      # element tags, node tags
        allTags, coord, _ = gmsh.model.mesh.getNodes()
        xyz = zeros(3, problem.non)
        xyz[:, allTags] .= reshape(coord, 3, :)
        conn = Vector{Int64}(nodeTags)                                  # 1-based, element-major
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:llfem_mesh_create, lib), Cint,
                    (Ptr{Cvoid}, Ptr{Float64}, Int64, Int32, Int32, Ptr{Int64}, Int64, Int32, Int32,
                     Ref{Ptr{Cvoid}}),
                    ctx(), xyz, problem.non, et, nn, conn, length(elemTags), problem.pdim, 1, r))
        etags = Vector{Int64}(elemTags)
        TypeMesh(Int32(et), Int32(nn), order, r[], etags, Dict(t => i for (i, t) in enumerate(etags)))
    end
end

# the volume element types of the problem, in gmsh's order
volume_types(problem::Problem) =
    (gmsh.model.setCurrent(problem.name);
     unique(Int32.(gmsh.model.mesh.getElementTypes(problem.dim, -1))))

function pattern(m::TypeMesh, op, ndof)
    nnz = Ref{Int64}(0)
    check(ccall((:llfem_mesh_pattern, lib), Cint, (Ptr{Cvoid}, Int32, Int32, Ref{Int64}, Ptr{Int64}, Ptr{Int64}),
                m.handle, op, 1, nnz, C_NULL, C_NULL))
    colptr = Vector{Int64}(undef, ndof + 1); rowval = Vector{Int64}(undef, nnz[])
    check(ccall((:llfem_mesh_pattern, lib), Cint, (Ptr{Cvoid}, Int32, Int32, Ref{Int64}, Ptr{Int64}, Ptr{Int64}),
                m.handle, op, 1, nnz, colptr, rowval))
    colptr, rowval
end

# 1.0 on the elements of material group `phName` that belong to this type,
# by element tag (the reference takes the group's elements per entity with
# getElementsByType(et, etag), src/linear.jl:261-263)
function group_scale(m::TypeMesh, phName)
    scale = zeros(length(m.elemTags))
    for (edim, etag) in gmsh.model.getEntitiesForPhysicalName(phName)
        elemTags, _ = gmsh.model.mesh.getElementsByType(m.et, etag)
        for t in elemTags
            scale[m.pos[Int64(t)]] = 1.0
        end
    end
    scale
end

# One operator on one element type: a llfem_assemble call per material group
# that has elements of this type, the first replacing and the others adding
# (accumulate) on the device, so the mesh holds the sum in group order;
# the last call copies the values out.  Returns the type's CSC (or diagonal).
function assemble_type(problem::Problem, m::TypeMesh, op::Int32; lumped=false, D=nothing)
    pts, w = gmsh.model.mesh.getIntegrationPoints(m.et, "Gauss" * string(2m.order + 1))
    _, dN, _ = gmsh.model.mesh.getBasisFunctions(m.et, pts, "GradLagrange")
    _, N, _ = gmsh.model.mesh.getBasisFunctions(m.et, pts, "Lagrange")
    ndof = problem.pdim * problem.non
    Dm = D === nothing ? nothing : Matrix{Float64}(permutedims(D))   # row-major 6x6
    lump = op == OP_MASS && lumped
    colptr, rowval = lump ? (Int64[], Int64[]) : pattern(m, op, ndof)
    nzval = zeros(lump ? ndof : length(rowval))
    groups = [(mat, group_scale(m, mat.phName)) for mat in problem.material]
    groups = filter(g -> any(!=(0.0), g[2]), groups)
    isempty(groups) && return lump ? nzval : SparseMatrixCSC(ndof, ndof, colptr, rowval, nzval)
    for (i, (mat, scale)) in enumerate(groups)
        whole = length(groups) == 1 && all(==(1.0), scale)
        acc = !lump && i > 1                 # lumped mass: the groups' diagonals are added here
        out = lump ? zeros(ndof) : nzval
        copy_out = lump || i == length(groups)
        GC.@preserve w N dN Dm scale out begin
            desc = OpDesc(op, length(w), pointer(w), pointer(N), pointer(dN),
                          Dm === nothing ? C_NULL : pointer(Dm), mat.E, mat.ν, mat.ρ, lumped, 0,
                          whole ? C_NULL : pointer(scale), acc)
            check(ccall((:llfem_assemble, lib), Cint, (Ptr{Cvoid}, Ref{OpDesc}, Ptr{Float64}, Ref{Stats}),
                        m.handle, desc, copy_out ? pointer(out) : C_NULL, Stats()))
        end
        lump && (nzval .+= out)
    end
    lump ? nzval : SparseMatrixCSC(ndof, ndof, colptr, rowval, nzval)
end

# stiffness / mass of the problem: the element types' operators added in
# type order (the reference appends every type's triplets before sparse(),
# src/linear.jl:225-321), then dropzeros! (src/linear.jl:324)
function assemble(problem::Problem, op::Int32; lumped=false, D=nothing)
    ndof = problem.pdim * problem.non
    parts = [assemble_type(problem, type_mesh(problem, et), op; lumped=lumped, D=D)
             for et in volume_types(problem)]
    if op == OP_MASS && lumped
        return SystemMatrix(dropzeros!(spdiagm(0 => reduce(+, parts))), problem)
    end
    SystemMatrix(dropzeros!(reduce(+, parts)), problem)
end

# many single-material, single-type problems in one launch: [(problem, op)];
# values stay on the device (read back per mesh with `values`)
function assemble_batch(items::Vector{Tuple{Problem,Int32}})
    descs = OpDesc[]; meshes = Ptr{Cvoid}[]; keep = Any[]
    for (problem, op) in items
        mat = problem.material[1]
        m = type_mesh(problem, only(volume_types(problem)))
        pts, w = gmsh.model.mesh.getIntegrationPoints(m.et, "Gauss" * string(2m.order + 1))
        _, dN, _ = gmsh.model.mesh.getBasisFunctions(m.et, pts, "GradLagrange")
        _, N, _ = gmsh.model.mesh.getBasisFunctions(m.et, pts, "Lagrange")
        push!(keep, (w, N, dN))
        push!(meshes, m.handle)
        push!(descs, OpDesc(op, length(w), pointer(w), pointer(N), pointer(dN), C_NULL,
                            mat.E, mat.ν, mat.ρ, 0, 0, C_NULL, 0))
    end
    GC.@preserve keep check(ccall((:llfem_assemble_batch, lib), Cint,
                                  (Int32, Ptr{Ptr{Cvoid}}, Ptr{OpDesc}, Ref{Stats}),
                                  length(descs), meshes, descs, Stats()))
    meshes
end

device_values(m::Ptr{Cvoid}, op, n) = (v = Vector{Float64}(undef, n);
                                       check(ccall((:llfem_mesh_values, lib), Cint, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}),
                                                   m, op, 0, v)); v)

# serving: one server per process; submit returns an id, wait blocks for the result
const SERVER = Ref{Ptr{Cvoid}}(C_NULL)
function server(; budget=0, max_active=64)
    if SERVER[] == C_NULL
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:llfem_server_create, lib), Cint, (Cint, Csize_t, Int32, Ref{Ptr{Cvoid}}),
                    0, budget, max_active, r))
        SERVER[] = r[]
    end
    SERVER[]
end
submit(req::Request) = (id = Ref{Int64}(0);
                        check(ccall((:llfem_server_submit, lib), Cint, (Ptr{Cvoid}, Ref{Request}, Ref{Int64}),
                                    server(), req, id)); id[])
wait_request(id) = (st = RequestStats();
                    check(ccall((:llfem_server_wait, lib), Cint, (Ptr{Cvoid}, Int64, Ref{RequestStats}),
                                server(), id, st)); st)

# the drop-in methods (dispatched by the extension on a backend flag)
gpu_stiffnessMatrix(problem) = assemble(problem, OP_STIFFNESS)
gpu_massMatrix(problem; lumped=true) = assemble(problem, OP_MASS; lumped=lumped)

# body force on the elements elems (0-based indices into the type mesh's
# elements): components numbers or functions of (x, y, z), evaluated at the
# volume Gauss points the library returns
function body_force(m::TypeMesh, elems::Vector{Int64}, comps, ndof)
    pts, w = gmsh.model.mesh.getIntegrationPoints(m.et, "Gauss" * string(2m.order + 1))
    _, dN, _ = gmsh.model.mesh.getBasisFunctions(m.et, pts, "GradLagrange")
    _, N, _ = gmsh.model.mesh.getBasisFunctions(m.et, pts, "Lagrange")
    q = length(w); ne = length(elems)
    fe = zeros(ndof)
    if any(c -> c isa Function, comps)
        xyz = zeros(3, q, ne)
        check(ccall((:llfem_volume_points, lib), Cint,
                    (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Int64}, Int64, Ptr{Float64}),
                    m.handle, q, N, elems, ne, xyz))
        g = zeros(3, q, ne)
        for e in 1:ne, j in 1:q, c in 1:3
            v = comps[c]
            g[c, j, e] = v isa Function ? v(xyz[1, j, e], xyz[2, j, e], xyz[3, j, e]) : v
        end
        check(ccall((:llfem_load_body, lib), Cint,
                    (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Int64}, Int64, Ptr{Float64},
                     Ptr{Float64}, Ptr{Float64}),
                    m.handle, q, w, N, dN, elems, ne, C_NULL, g, fe))
    else
        check(ccall((:llfem_load_body, lib), Cint,
                    (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Int64}, Int64,
                     Ref{NTuple{3,Float64}}, Ptr{Float64}, Ptr{Float64}),
                    m.handle, q, w, N, dN, elems, ne, Float64.(comps), C_NULL, fe))
    end
    fe
end

# loadVector(problem, loads) for a 3D solid (src/linear.jl:3088-3434): on a
# surface group, a traction (fx, fy, fz) or a pressure p (src/linear.jl:3196-3224;
# p replaces fx, fy, fz); on a curve group a line load and on a point group a
# concentrated force (the library takes |dx/dxi| or 1 as the measure by the
# element type, src/linear.jl:3385-3388); on a volume group, a body force.  Components and p
# are numbers or functions of (x, y, z): traction and body-force functions
# are evaluated at the Gauss points the library returns, p at the nodes
# (scalarField, src/general.jl:4477-4515).
function gpu_loadVector(problem::Problem, loads)
    loads isa Vector || error("loadVector: loads must be in a vector.")
    ets = volume_types(problem)
    m = type_mesh(problem, first(ets))                 # any type's mesh holds every node
    ndof = problem.non * problem.pdim
    f = zeros(ndof)
    for ld in loads
        (ld.problem === nothing || ld.problem === problem) || continue
        val(k) = (x = get(ld.values, k, nothing); x === nothing ? 0.0 : x)
        comps = (val(:fx), val(:fy), val(:fz))
        p = get(ld.values, :p, nothing)
        for (edim, etag) in gmsh.model.getEntitiesForPhysicalName(ld.phName)
            if edim == 3
                p === nothing || error("loadVector: pressure can be given on a surface of a 3D solid.")
                for et in ets
                    tm = type_mesh(problem, et)
                    elemTags, _ = gmsh.model.mesh.getElementsByType(tm.et, etag)
                    isempty(elemTags) && continue
                    elems = Int64[tm.pos[Int64(t)] - 1 for t in elemTags]
                    f .+= body_force(tm, elems, comps, ndof)
                end
                continue
            end
            p === nothing || edim == 2 || error("loadVector: pressure can be given on a surface of a 3D solid.")
            ets2, _, ents = gmsh.model.mesh.getElements(edim, etag)
            for (et, enodes) in zip(ets2, ents)
                _, _, order, nf, lnc, _ = gmsh.model.mesh.getElementProperties(et)
                pts, w = gmsh.model.mesh.getIntegrationPoints(et, "Gauss" * string(2order + 1))
                _, Nr, _ = gmsh.model.mesh.getBasisFunctions(et, pts, "Lagrange")
                _, dN, _ = gmsh.model.mesh.getBasisFunctions(et, pts, "GradLagrange")
                F = length(enodes) ÷ nf; q = length(w)
                conn = Vector{Int64}(enodes)            # 1-based tags, element-major
                fe = zeros(ndof)
                if p !== nothing
                    # basis gradients at the face's own nodes, as normalVector takes them
                    lnc3 = zeros(3nf); lnc3[1:3:end] .= lnc[1:2:end]; lnc3[2:3:end] .= lnc[2:2:end]
                    _, dNn, _ = gmsh.model.mesh.getBasisFunctions(et, lnc3, "GradLagrange")
                    pn = C_NULL
                    if p isa Function
                        tags, xyz, _ = gmsh.model.mesh.getNodes(edim, etag, true, false)
                        pv = zeros(problem.non)
                        pv[tags] .= p.(xyz[1:3:end], xyz[2:3:end], xyz[3:3:end])
                        pn = pv
                    end
                    check(ccall((:llfem_load_pressure, lib), Cint,
                                (Ptr{Cvoid}, Int32, Int32, Ptr{Int64}, Int64, Int32, Int32, Ptr{Float64},
                                 Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Ptr{Float64}, Ptr{Float64}),
                                m.handle, et, nf, conn, F, 1, q, w, Nr, dN, dNn,
                                p isa Function ? 0.0 : Float64(p), pn, fe))
                elseif any(c -> c isa Function, comps)
                    xyz = zeros(3, q, F)
                    check(ccall((:llfem_face_points, lib), Cint,
                                (Ptr{Cvoid}, Int32, Int32, Ptr{Int64}, Int64, Int32, Int32, Ptr{Float64},
                                 Ptr{Float64}), m.handle, et, nf, conn, F, 1, q, Nr, xyz))
                    t = zeros(3, q, F)
                    for e in 1:F, j in 1:q, c in 1:3
                        v = comps[c]
                        t[c, j, e] = v isa Function ? v(xyz[1, j, e], xyz[2, j, e], xyz[3, j, e]) : v
                    end
                    check(ccall((:llfem_load_traction_gp, lib), Cint,
                                (Ptr{Cvoid}, Int32, Int32, Ptr{Int64}, Int64, Int32, Int32, Ptr{Float64},
                                 Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                                m.handle, et, nf, conn, F, 1, q, w, Nr, dN, t, fe))
                else
                    check(ccall((:llfem_load_traction, lib), Cint,
                                (Ptr{Cvoid}, Int32, Int32, Ptr{Int64}, Int64, Int32, Int32, Ptr{Float64},
                                 Ptr{Float64}, Ptr{Float64}, Ref{NTuple{3,Float64}}, Ptr{Float64}),
                                m.handle, et, nf, conn, F, 1, q, w, Nr, dN, Float64.(comps), fe))
                end
                f .+= fe
            end
        end
    end
    VectorField([], reshape(f, :, 1), [0.0], [], 1, :v3D, problem)
end

# solveDisplacement(K, f; support, iterative=false, reltol, maxiter, preconditioner)
# (src/linear.jl:3711-3743).  iterative=false is the reference's direct solve
# K[free,free] \ rhs: here the same system is solved by FP64 PCG to a relative
# residual of 1e-12; iterative=true honours reltol and maxiter.  The
# reference's Pl=preconditioner maps to Jacobi, or to the AMG V-cycle for
# preconditioner = :amg (single element type: the stiffness kept on the
# device mesh by gpu_stiffnessMatrix is solved).
function gpu_solveDisplacement(K::SystemMatrix, f::VectorField; support=BoundaryCondition[],
                               iterative=false, reltol=sqrt(eps()), maxiter=K.model.non * K.model.dim,
                               preconditioner=nothing, ordering=true, mixed=false)
    f.type == :v3D || error("solveDisplacement: wrong type of 'f': ($(f.type))")
    problem = K.model
    fixed = Vector{Int64}(LowLevelFEM.constrainedDoFs(problem, support))   # 1-based
    u0 = copy(f); fill!(u0.a, 0.0); LowLevelFEM.applyBoundaryConditions!(u0, support)
    ufix = u0.a[fixed, 1]
    tol = iterative ? Float64(reltol) : 1e-12
    its = iterative ? Int64(maxiter) : Int64(max(maxiter, 10 * size(K.A, 1)))
    rhs = Vector{Float64}(f.a[:, 1])
    u = copy(f)
    sol = zeros(length(rhs))
    if preconditioner === :amg
        ets = volume_types(problem)
        length(ets) == 1 || error("solveDisplacement: preconditioner = :amg needs a single element type")
        m = MESHES[(problem.name, only(ets))]
        check(ccall((:llfem_solve, lib), Cint,
                    (Ptr{Cvoid}, Ptr{Float64}, Ptr{Int64}, Int64, Int32, Ptr{Float64}, Int32, Float64,
                     Int64, Ptr{Float64}, Ref{Stats}),
                    m.handle, rhs, fixed, length(fixed), 1, ufix, PCG_AMG_RELIABLE, tol, its, sol, Stats()))
    else
        A = K.A
        colptr = Vector{Int64}(A.colptr); rowval = Vector{Int64}(A.rowval)
        check(ccall((:llfem_solve_csc, lib), Cint,
                    (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Int32, Ptr{Float64},
                     Ptr{Int64}, Int64, Ptr{Float64}, Int32, Float64, Int64, Ptr{Float64}, Ref{Stats}),
                    ctx(), size(A, 1), colptr, rowval, A.nzval, 1, rhs, fixed, length(fixed), ufix,
                    mixed ? PCG_MIXED_REFINE : PCG_JACOBI_FP64, tol, its, sol, Stats()))
    end
    u.a[:, 1] .= sol
    u
end

end # module
