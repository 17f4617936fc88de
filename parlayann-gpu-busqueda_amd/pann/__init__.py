"""pann: ParlayANN graph beam search on MI355X (gfx950), behind the C ABI of include/pann.h.

Modules:
    search    host mirror of VectoresGPU / GrafoGPU / HnswGPU / BuscadorGPU / BuscadorHnsw
    formats   .i8bin, ParlayANN graph, HNSW v3, ground truth and neighbour files
    datagen   synthetic Gaussian-mixture int8 data
    builders  CPU graph builders (Vamana, HNSW, HCNNG, PyNNDescent)
    recall    tie-aware recall@k
    report    neighbour-by-neighbour validation and recall-QPS tables (python -m pann.report)
"""
from . import datagen, formats, recall  # noqa: F401
from ._lib import PannError  # noqa: F401
