"""alamo_amd: MI355X-native nodal elastic multigrid (Alamo's Operator::Elastic /
Solver::Nonlocal::Linear path) behind the C-ABI of include/alamo_elastic.h."""
from . import bc, model, problems  # noqa: F401

__all__ = ["bc", "model", "problems", "elastic", "build"]
