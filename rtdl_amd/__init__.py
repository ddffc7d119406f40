"""rtdl_amd — an MI355X-native backend for the RTDL kernel contract.

Use it as the reference's ``rtdsl`` module is used::

    import rtdl_amd as rt

    @rt.kernel(backend="rtdl", precision="float_approx")
    def frn():
        q = rt.input("query_points", rt.Points3D, layout=rt.Point3DLayout, role="probe")
        s = rt.input("search_points", rt.Points3D, layout=rt.Point3DLayout, role="build")
        c = rt.traverse(q, s, accel="bvh")
        h = rt.refine(c, predicate=rt.fixed_radius_neighbors(radius=0.02, k_max=64))
        return rt.emit(h, fields=["query_id", "neighbor_id", "distance"])

    rows = rt.run_mi355x(frn, query_points=..., search_points=...)
    result = rt.run(frn, inputs, execution=rt.ExecutionPolicy(backend="mi355x"))
"""
from .api import (
    bounded_knn_rows,
    compile_kernel,
    contains,
    emit,
    fixed_radius_neighbors,
    input,
    kernel,
    knn_rows,
    point_in_polygon,
    ray_triangle_any_hit,
    ray_triangle_closest_hit,
    ray_triangle_hit_count,
    refine,
    traverse,
)
from .execution import (
    ExecutionPolicy,
    ExecutionReport,
    ExecutionResult,
    register_backend,
    registered_backends,
    run,
)
from .ir import (
    CandidateSet,
    CompiledKernel,
    EmitOp,
    GeometryInput,
    Predicate,
    RefineOp,
    RTExecutionPlan,
)
from .layout_types import (
    Point2DLayout,
    Point3DLayout,
    Points,
    Points3D,
    Polygon2DLayout,
    Polygons,
    Ray2DLayout,
    Ray3DLayout,
    Rays,
    Rays3D,
    Triangle2DLayout,
    Triangle3DLayout,
    Triangles,
    Triangles3D,
)
from .lowering import lower_to_execution_plan, lower_to_rayjoin
from .mi355x_runtime import (
    PreparedMi355xKernel,
    mi355x_available,
    mi355x_context_probe,
    mi355x_version,
    prepare_mi355x,
    run_mi355x,
)
from .records import Point, Point3D, Polygon, Ray2D, Ray3D, Triangle, Triangle3D
from .visibility import visibility_pair_rows, visibility_rows

__all__ = [name for name in dir() if not name.startswith("_")]
