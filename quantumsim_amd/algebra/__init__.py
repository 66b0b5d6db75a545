from .algebra import kraus_to_ptm, ptm_convert_basis, dm_to_pv, pv_to_dm

__all__ = ['kraus_to_ptm', 'ptm_convert_basis', 'dm_to_pv', 'pv_to_dm']
