"""Physical constants (CODATA 2018 / SI 2019), the values
ion_gym/physics/constants.py:18-37 defines, so that every derived scalar
(accelerations, gas speeds, SDS numbers) is the reference's number."""
E_CHG = 1.602176634e-19        # C, exact
K_B = 1.380649e-23             # J/K, exact
N_A = 6.02214076e23            # 1/mol, exact
KG_AMU = 1.66053906660e-27     # kg per Da, CODATA 2018
T_STP_K = 273.15
P_STP_PA = 101325.0
ATM_TORR = 760.0
TORR_PA = P_STP_PA / ATM_TORR
K_B_EV = K_B / E_CHG
ACC = E_CHG / KG_AMU * 1e-6
