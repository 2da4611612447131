"""Python mirror of the reference's museqa.sequence / museqa.database /
museqa.pairwise modules (src/python/*.pyx) plus museqa.phylogeny and
museqa.pgalign, over the
C++ host layer (msa_amd/csrc/host) and the HIP kernels behind libmsa_hip.so."""
from . import database, pairwise, pgalign, phylogeny, sequence  # noqa: F401
