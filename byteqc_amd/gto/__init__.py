from .mole import Mole, M, BOHR, nf_cart, gto_norm, load_basis, ANGULAR
from .cart2sph import cart2sph_block, cart_components
from .clusters import water_cluster, water_cluster_coords, water_monomer
