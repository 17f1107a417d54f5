# SANKHYA solution file
# generated-by sih26_amd libsk_gpu (MI355X PDHG engine)
# All numbers carry 17 significant digits and round-trip exactly.
model closed_equality
source 
sense minimize
status optimal
algorithm pdhg-cpu
objective 4.0000000000000018
dual_bound 3.9999999999999969
mip_relative_gap 0.0001
mip_absolute_gap 9.9999999999999995e-07
objective_offset 0
certificate none
rows 1
columns 2
iterations 77
nodes 0
solve_seconds 0
primal_infeasibility 1.7763568394002505e-15
dual_infeasibility 0
integrality_violation 0

# name value reduced_cost basis_status
begin columns 2
C0 4.0000000000000018 7.7715611723760958e-16 unknown
C1 0 1.0000000000000009 unknown
end columns

# name activity dual basis_status
begin rows 1
R0 4.0000000000000018 0.99999999999999922 unknown
end rows
